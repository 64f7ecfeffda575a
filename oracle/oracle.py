"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the CPU oracle (oracle/_build/libepsx_oracle.so,
built from epsx_oracle.c by oracle/Makefile).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module;
the product (epspectral.jl_amd/) never does.  Functions take plain numpy arrays in the
reference's column-major index order, or a setup object with the attributes of the reference's
`setup` NamedTuple (g2, ωph, iq_kk, iks_i, iks_f, ε_i, w_i, ε_f, w_f, f_to_i, W, nstates_base,
multiplets, ωs, ωs_spectral, μ_start), as the tests build it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libepsx_oracle.so")

P, I64, I32, D = C.c_void_p, C.c_int64, C.c_int32, C.c_double


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sig = {
            "orc_occ_fermion": (D, [D, D]),
            "orc_occ_boson": (D, [D, D]),
            "orc_grid_from_nodes": (C.c_int, [I64, P, D, P, P, P, P, P, P]),
            "orc_graded_grid": (I64, [I64, P, P, P, P, I64]),
            "orc_cell_search_table": (I64, [I64, P, D, P, P, P, I64]),
            "orc_interp_linear_grid": (None, [P, I64, D, I64, P, P, D, D, P]),
            "orc_linear_interpolation_table": (None, [I64, P, I64, P, P, P]),
            "orc_apply_interpolation_table": (None, [I64, I64, I64, P, P, P, P]),
            "orc_trapezoid_weights": (None, [I64, P, P]),
            "orc_kramers_kronig_matrix": (None, [I64, P, P]),
            "orc_fm_term": (D, [D, D, D, D, I64, P, I64, P, P, D, D, D, D, D, D]),
            "orc_add_fan_migdal": (None, [I64, I64, I64, I64, I64, P, P, P, P, P, I64, P, P, D, D,
                                          I64, P, D, D, D, P, C.c_int]),
            "orc_loop_scgd0": (C.c_int, [P, P, P]),
            "orc_ncarrier_spectral": (D, [D, D, I64, P, P, D, C.c_int, D, D]),
            "orc_mu_bisection": (D, [D, D, D, I64, P, P, D, C.c_int, D, D, D, P]),
            "orc_build_multiplets": (I64, [I64, P, P, D, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c(fn, *args):
    """Call a foreign function with numpy arrays passed as data pointers.  The arrays stay
    referenced by `args` until the call returns (a pointer taken from a temporary copy would
    dangle once the copy is freed)."""
    return fn(*[a.ctypes.data if isinstance(a, np.ndarray) else a for a in args])


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def _i32(a):
    return np.asfortranarray(a, dtype=np.int32)


def _i64(a):
    return np.asfortranarray(a, dtype=np.int64)


def occ_fermion(e, T):
    return _c(lib().orc_occ_fermion, float(e), float(T))


def occ_boson(e, T):
    return _c(lib().orc_occ_boson, float(e), float(T))


class OracleError(Exception):
    pass


def grid_from_nodes(ws, atol=1e-8):
    """(origin, unit, offsets) or raises OracleError(code, bad_j, dev, best_unit)."""
    ws = _f64(ws)
    n = len(ws)
    o, u, dev, bu = np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1)
    j = np.zeros(1, dtype=np.int64)
    off = np.zeros(n, dtype=np.int64)
    rc = _c(lib().orc_grid_from_nodes, n, (ws), atol, (o), (u), (off), (dev), (j), (bu))
    if rc:
        raise OracleError(rc, int(j[0]), float(dev[0]), float(bu[0]))
    return float(o[0]), float(u[0]), off


def graded_grid(windows):
    w = _f64(np.asarray(windows, dtype=np.float64).ravel())
    cap = 1 << 22
    off = np.zeros(cap, dtype=np.int64)
    o, u = np.zeros(1), np.zeros(1)
    n = _c(lib().orc_graded_grid, len(windows), (w), (o), (u), (off), cap)
    if n < 0:
        raise OracleError(int(n))
    return float(o[0]), float(u[0]), off[:n].copy()


def cell_search_table(ws, dw_min):
    ws = _f64(ws)
    x0, ih = np.zeros(1), np.zeros(1)
    n = _c(lib().orc_cell_search_table, len(ws), (ws), dw_min, (x0), (ih), None, 0)
    if n < 0:
        raise OracleError(int(n))
    jtab = np.zeros(n, dtype=np.int32)
    _c(lib().orc_cell_search_table, len(ws), (ws), dw_min, (x0), (ih), (jtab), n)
    return jtab, float(x0[0]), float(ih[0])


def interp_linear_grid(ys, col, x, ws, table):
    ys = np.asfortranarray(ys, dtype=np.complex128)
    ws = _f64(ws)
    jtab, x0, ih = table
    out = np.zeros(2)
    _c(lib().orc_interp_linear_grid, (ys), col, float(x), len(ws), (ws), (jtab), x0, ih, (out))
    return complex(out[0], out[1])


def linear_interpolation_table(ws, xs):
    ws, xs = _f64(ws), _f64(xs)
    j = np.zeros(len(xs), dtype=np.int32)
    t = np.zeros(len(xs))
    _c(lib().orc_linear_interpolation_table, len(ws), (ws), len(xs), (xs), (j), (t))
    return j, t


def apply_interpolation_table(ys, j, t):
    ys = np.asfortranarray(ys, dtype=np.complex128)
    ys2 = ys.reshape(ys.shape[0], -1, order="F")
    Y = np.zeros((len(j), ys2.shape[1]), dtype=np.complex128, order="F")
    _c(lib().orc_apply_interpolation_table, len(j), ys2.shape[0], ys2.shape[1], (ys2), (_i32(j)),
                                        (_f64(t)), (Y))
    return Y


def trapezoid_weights(ws):
    ws = _f64(ws)
    W = np.zeros(len(ws))
    _c(lib().orc_trapezoid_weights, len(ws), (ws), (W))
    return W


def kramers_kronig_matrix(ws):
    ws = _f64(ws)
    K = np.zeros((len(ws), len(ws)), order="F")
    _c(lib().orc_kramers_kronig_matrix, len(ws), (ws), (K))
    return K


def add_fan_migdal(ImS, g2, wq_fi, e_f, f_to_i, w_f, ws, table, Sin, mu, T, w_acoustic, nthreads=8,
                   f0=0, f1=None):
    """ImS[ω, i] += Σ_{f in [f0, f1), ν} term, with wq_fi the pair table [ν, f, i]."""
    nm, ni, nf = g2.shape
    f1 = nf if f1 is None else f1
    ImS = np.asfortranarray(ImS, dtype=np.float64)
    jtab, x0, ih = table
    _c(lib().orc_add_fan_migdal, nm, ni, nf, f0, f1, (_f64(g2)), (_f64(wq_fi)), (_f64(e_f)),
                             (_i32(f_to_i)), (_f64(w_f)), len(ws), (_f64(ws)), (jtab), x0, ih,
                             Sin.shape[1], (np.asfortranarray(Sin, dtype=np.complex128)), mu, T,
                             w_acoustic, (ImS), nthreads)
    return ImS


def ncarrier_spectral(mu, T, wsa, D, W, count_holes, nstates_base, spin_degeneracy):
    return _c(lib().orc_ncarrier_spectral, mu, T, len(wsa), (_f64(wsa)), (_f64(D)), W, int(count_holes),
                                       nstates_base, spin_degeneracy)


def build_multiplets(iks, es, cutoff):
    ni = len(es)
    ptr = np.zeros(ni + 1, dtype=np.int64)
    idx = np.zeros(ni, dtype=np.int64)
    nm = _c(lib().orc_build_multiplets, ni, (_i32(iks)), (_f64(es)), cutoff, (ptr), (idx))
    return [list(idx[ptr[m]:ptr[m + 1]]) for m in range(nm)]


class _Setup(C.Structure):
    _fields_ = [("nm", I64), ("ni", I64), ("nf", I64), ("nw", I64), ("nwa", I64), ("nq", I64),
                ("nki", I64), ("nkf", I64),
                ("g2", P), ("wph", P), ("iq_kk", P), ("iks_i", P), ("iks_f", P),
                ("e_i", P), ("w_i", P), ("e_f", P), ("w_f", P), ("f_to_i", P),
                ("W", D), ("nstates_base", D), ("ws", P), ("dw_min", D), ("wsa", P),
                ("n_mult", I64), ("mult_ptr", P), ("mult_idx", P)]


class _Params(C.Structure):
    _fields_ = [("nT", I64), ("Tlist", P), ("nlist", P), ("mu_start", P), ("eta_min", P),
                ("nelec", D), ("spin_degeneracy", D), ("is_semiconductor", C.c_int),
                ("eta_init", D), ("tol", D), ("w_acoustic", D), ("maxiter", I64), ("nf_chunk", I64),
                ("fix_mu", C.c_int), ("Sigma_init", P), ("nthreads", C.c_int)]


class _Out(C.Structure):
    _fields_ = [("Sigma", P), ("mu", P), ("converged", P), ("hist_err", P), ("hist_mu", P),
                ("hist_minA", P), ("wall", P), ("iter_frozen", P), ("niter", I64), ("status", C.c_int),
                ("err_T", I64), ("err_n_target", D), ("err_bracket", D * 2)]


def loop_scgd0(setup, Tlist, nlist, nelec, *, η_min, η_init, maxiter, tol, ω_acoustic,
               spin_degeneracy=2.0, semiconductor=False, fix_μ=False, μ_start=None, Σ_init=None,
               gpu_tile=None, nthreads=8):
    """`_loop_scgd0` (src/scgd0.jl:189-346) on the CPU.  Returns dict(Σ, μ, converged, history)
    or raises OracleError(2, T index) when no μ root is bracketed."""
    s = setup
    nm, ni, nf = s.g2.shape
    keep = dict(g2=_f64(s.g2), wph=_f64(s.ωph), iq=_i32(s.iq_kk), iki=_i32(s.iks_i), ikf=_i32(s.iks_f),
                ei=_f64(s.ε_i), wi=_f64(s.w_i), ef=_f64(s.ε_f), wf=_f64(s.w_f), fi=_i32(s.f_to_i),
                ws=_f64(np.asarray(s.ωs)), wsa=_f64(np.asarray(s.ωs_spectral)))
    mptr, midx = [0], []
    for m in s.multiplets:
        midx.extend(m)
        mptr.append(len(midx))
    keep["mptr"] = _i64(mptr)
    keep["midx"] = _i64(midx if midx else [0])
    S = _Setup(nm, ni, nf, len(keep["ws"]), len(keep["wsa"]), keep["wph"].shape[1], keep["iq"].shape[0],
               keep["iq"].shape[1], _p(keep["g2"]), _p(keep["wph"]), _p(keep["iq"]), _p(keep["iki"]),
               _p(keep["ikf"]), _p(keep["ei"]), _p(keep["wi"]), _p(keep["ef"]), _p(keep["wf"]),
               _p(keep["fi"]), float(s.W), float(s.nstates_base), _p(keep["ws"]), float(s.ωs.dω_min),
               _p(keep["wsa"]), len(s.multiplets), _p(keep["mptr"]), _p(keep["midx"]))
    Tl = _f64(np.atleast_1d(Tlist))
    nT = len(Tl)
    nl = _f64(np.broadcast_to(np.asarray(nlist, dtype=float), (nT,)))
    mu0 = _f64(np.broadcast_to(np.asarray(s.μ_start if μ_start is None else μ_start, dtype=float), (nT,)))
    eta = _f64(np.broadcast_to(np.asarray(η_min, dtype=float), (nT,)))
    Si = None if Σ_init is None else np.asfortranarray(Σ_init, dtype=np.complex128)
    Pm = _Params(nT, _p(Tl), _p(nl), _p(mu0), _p(eta), float(nelec), float(spin_degeneracy),
                 int(semiconductor), float(η_init), float(tol), float(ω_acoustic), int(maxiter),
                 0 if gpu_tile is None else int(gpu_tile), int(fix_μ), _p(Si), int(nthreads))
    nw = len(keep["ws"])
    Σ = np.zeros((nw, ni, nT), dtype=np.complex128, order="F")
    μ = np.zeros(nT)
    conv = np.zeros(nT, dtype=np.int32)
    he, hm, ha = (np.zeros((nT, maxiter), order="F") for _ in range(3))
    wall = np.zeros(maxiter)
    fr = np.zeros(nT, dtype=np.int64)
    O = _Out(_p(Σ), _p(μ), _p(conv), _p(he), _p(hm), _p(ha), _p(wall), _p(fr), 0, 0, 0, 0.0, (D * 2)(0, 0))
    rc = _c(lib().orc_loop_scgd0, C.byref(S), C.byref(Pm), C.byref(O))
    if rc:
        raise OracleError(rc, int(O.err_T))
    n = int(O.niter)
    return dict(Σ=Σ, μ=μ, converged=conv.astype(bool),
                history=dict(err=he[:, :n].copy(), μ=hm[:, :n].copy(), min_spectral_weight=ha[:, :n].copy(),
                             wall=wall[:n].copy(), iter_frozen=fr.copy()))


# ------------------------------------------------------------------------------ ladder
def _ladder_sigs():
    L = lib()
    if getattr(L, "_ladder_ready", False):
        return L
    sig = {
        "orc_contour_green": (None, [D, D, D, D, D, D, C.c_int, P]),
        "orc_contour_greens": (None, [I64, I64, P, P, P, P, D, D, D, P, P]),
        "orc_attach_greens": (None, [I64, I64, P, P, P, P]),
        "orc_attach_greens_bare": (None, [I64, I64, P, P, P, P]),
        "orc_conductivity_from_vertex": (None, [I64, I64, P, P, P, P, D, D, P]),
        "orc_vertex_shift_tables": (None, [I64, P, I64, P, P, P]),
        "orc_apply_ladder_kernel": (C.c_int, [P, P, P, P, D, D, I64, P, C.c_int]),
        "orc_ladder_solve": (C.c_int, [P, P, P, P, P, P, D, D, D, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._ladder_ready = True
    return L


def _c128(a):
    return np.asfortranarray(a, dtype=np.complex128)


def contour_green(w, e, S, mu, T, c12):
    out = np.zeros(2)
    _c(_ladder_sigs().orc_contour_green, w, e, S.real, S.imag, mu, T, c12, (out))
    return complex(out[0], out[1])


def contour_greens(ws, e_i, Sd_m, Sd_p, Omega, mu, T):
    ws = _f64(ws)
    nw, ni = len(ws), len(e_i)
    Gm = np.zeros((nw, 4, ni), dtype=np.complex128, order="F")
    Gp = np.zeros_like(Gm)
    _c(_ladder_sigs().orc_contour_greens, nw, ni, (ws), (_f64(e_i)), (_c128(Sd_m)), (_c128(Sd_p)), Omega, mu, T,
                                      (Gm), (Gp))
    return Gm, Gp


def attach_greens(Gm, X, Gp):
    nw, _, ni = Gm.shape
    J = np.zeros((nw, 4, 3, 2, ni), dtype=np.complex128, order="F")
    _c(_ladder_sigs().orc_attach_greens, nw, ni, (_c128(Gm)), (_c128(X)), (_c128(Gp)), (J))
    return J


def attach_greens_bare(Gm, v, Gp):
    nw, _, ni = Gm.shape
    J = np.zeros((nw, 4, 3, 2, ni), dtype=np.complex128, order="F")
    _c(_ladder_sigs().orc_attach_greens_bare, nw, ni, (_c128(Gm)), (_f64(v)), (_c128(Gp)), (J))
    return J


def conductivity_from_vertex(J, v, w_i, Ww, Omega, T):
    nw = J.shape[0]
    ni = J.shape[4]
    s = np.zeros((3, 3), order="F")
    _c(_ladder_sigs().orc_conductivity_from_vertex, nw, ni, (_c128(J)), (_f64(v)), (_f64(w_i)), (_f64(Ww)),
                                                Omega, T, (s))
    return s


def vertex_shift_tables(offsets, nshift):
    off = _i64(offsets)
    nw = len(off)
    klo = np.zeros((nw, nshift, 2), dtype=np.int32, order="F")
    wlo = np.zeros((nw, nshift, 2), order="F")
    whi = np.zeros((nw, nshift, 2), order="F")
    _c(_ladder_sigs().orc_vertex_shift_tables, nw, (off), nshift, (klo), (wlo), (whi))
    return klo, wlo, whi


class _LSetup(C.Structure):
    _fields_ = [("nm", I64), ("ni", I64), ("nf", I64), ("nw", I64), ("nq", I64), ("nki", I64), ("nkf", I64),
                ("nshift", I64), ("g2", P), ("wph", P), ("iq_kk", P), ("iks_i", P), ("iks_f", P), ("f_to_i", P),
                ("w_f", P), ("R", P), ("ws", P), ("offsets", P), ("dw_unit", D), ("KK", P)]


class _LParams(C.Structure):
    _fields_ = [("solver", C.c_int), ("rtol", D), ("maxiter", I64), ("warm", C.c_int), ("gpu_tile", I64),
                ("w_acoustic", D), ("nthreads", C.c_int)]


class _LOut(C.Structure):
    _fields_ = [("converged", I32), ("niter", I64), ("n_norms", I64), ("norms", P), ("sigma_iter", P),
                ("sigma_bubble", D * 9), ("sigma_ladder", D * 9)]


def _lsetup(ld):
    """ld: an object with the fields of the reference's ladder setup (g2, ωph, iq_kk, iks_i, iks_f,
    f_to_i, w_f, R, ωs (FrequencyGrid), nshift, KK)."""
    nm, ni, nf = ld.g2.shape
    ws = np.asarray(ld.ωs)
    keep = dict(g2=_f64(ld.g2), wph=_f64(ld.ωph), iq=_i32(ld.iq_kk), iki=_i32(ld.iks_i), ikf=_i32(ld.iks_f),
                fi=_i32(ld.f_to_i), wf=_f64(ld.w_f), R=_f64(ld.R), ws=_f64(ws), off=_i64(ld.ωs.offsets),
                KK=_f64(ld.KK if getattr(ld, "KK", None) is not None else kramers_kronig_matrix(ws)))
    S = _LSetup(nm, ni, nf, len(ws), keep["wph"].shape[1], keep["iq"].shape[0], keep["iq"].shape[1], int(ld.nshift),
                _p(keep["g2"]), _p(keep["wph"]), _p(keep["iq"]), _p(keep["iki"]), _p(keep["ikf"]), _p(keep["fi"]),
                _p(keep["wf"]), _p(keep["R"]), _p(keep["ws"]), _p(keep["off"]), float(ld.ωs.dω_unit), _p(keep["KK"]))
    return S, keep


def apply_ladder_kernel(ld, J, Gm, Gp, T, w_acoustic, gpu_tile=None, nthreads=8):
    S, keep = _lsetup(ld)
    LJ = np.zeros(J.shape, dtype=np.complex128, order="F")
    _c(_ladder_sigs().orc_apply_ladder_kernel, C.byref(S), (_c128(J)), (_c128(Gm)), (_c128(Gp)), T, w_acoustic,
                                           0 if gpu_tile is None else int(gpu_tile), (LJ), nthreads)
    return LJ


def ladder_solve(ld, Sigma_T, mu, T, Omega, *, solver="gmres", rtol=1e-2, maxiter=30, J=None, gpu_tile=None,
                 w_acoustic=1e-6, nthreads=8):
    """One (T, Ω) of `_loop_ladder` (src/ladder.jl:327-371).  J (optional) is the warm start;
    returns dict(J, σ_bubble, σ_ladder, converged, niter, norms, σ_iter)."""
    S, keep = _lsetup(ld)
    nw, ni = len(ld.ωs), ld.g2.shape[1]
    warm = J is not None
    Jw = _c128(J).copy(order="F") if warm else np.zeros((nw, 4, 3, 2, ni), dtype=np.complex128, order="F")
    norms = np.zeros(maxiter + 2)
    sit = np.zeros(9 * maxiter)
    O = _LOut(0, 0, 0, _p(norms), _p(sit), (D * 9)(), (D * 9)())
    Pm = _LParams(0 if solver == "gmres" else 1, float(rtol), int(maxiter), int(warm),
                  0 if gpu_tile is None else int(gpu_tile), float(w_acoustic), int(nthreads))
    Ww = trapezoid_weights(np.asarray(ld.ωs))
    _c(_ladder_sigs().orc_ladder_solve, C.byref(S), (_c128(Sigma_T)), (_f64(ld.ε_i)), (_f64(ld.w_i)), (_f64(ld.v)),
                                    (_f64(Ww)), float(Omega), float(mu), float(T), C.byref(Pm), (Jw), C.byref(O))
    return dict(J=Jw, σ_bubble=np.array(O.sigma_bubble[:]).reshape(3, 3, order="F"),
                σ_ladder=np.array(O.sigma_ladder[:]).reshape(3, 3, order="F"), converged=bool(O.converged),
                niter=int(O.niter), norms=norms[:O.n_norms].copy(),
                σ_iter=[sit[9 * k:9 * k + 9].reshape(3, 3, order="F").copy() for k in range(O.niter)] if solver != "gmres" else [])
