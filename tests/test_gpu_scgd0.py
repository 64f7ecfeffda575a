"""GPU parity of the scGD0 path: libepsx (through its C ABI) against the CPU oracle.

Tolerances: 1e-13 relative for single kernels in FP64 (the device's exp/log/division differ from
glibc's by an ulp), 1e-10 relative for whole scGD0 runs in FP64 (the reference's own GPU gate,
test/test_scgd0_gpu.jl:36-37), bitwise for run-to-run repeats and for integer tables.
"""
import math
import os
import warnings

import numpy as np
import pytest

import epsx
import oracle
from epsx.grids import FrequencyGrid, graded_frequency_grid
from epsx.models import holstein_1d_setup, t2g_setup, bare_mu
from ref_numpy import pseudorandom, pseudorandom_complex, naive_fan_migdal, holstein_bare_mu

pytestmark = pytest.mark.gpu

SRVO3_NODES = np.concatenate([np.arange(-350, -109, 20), np.arange(-99, 101), np.arange(120, 341, 20)]) / 1000.0  # Julia's correctly rounded ranges


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


def test_kk_matrix_and_apply(ctx):
    g = FrequencyGrid.from_nodes(SRVO3_NODES)
    K = epsx.kramers_kronig_matrix(g, ctx)
    K_ref = oracle.kramers_kronig_matrix(np.asarray(g))
    assert rel(K, K_ref) <= 1e-13
    z = pseudorandom(len(g), 5, seed=1)
    y = epsx.kramers_kronig(g, z, ctx)
    assert rel(y, K_ref @ z) <= 1e-13
    # a non-square shape through the MFMA GEMM edges
    g2 = FrequencyGrid(0.0, 0.01, np.arange(77))
    z2 = pseudorandom(77, 3, seed=2)
    assert rel(epsx.kramers_kronig(g2, z2, ctx), oracle.kramers_kronig_matrix(np.asarray(g2)) @ z2) <= 1e-13


def test_gemm_tile_shapes_give_the_same_bits(ctx):
    """Every FP64 GEMM tile shape runs the same k-tiles in the same order, so each gives the same
    bits: Kramers-Kronig on a 700-node grid applied to 300 columns (several
    tiles and k splits), each shape forced in turn, against the oracle at 1e-13."""
    g = FrequencyGrid(-1.0, 0.003, np.arange(700))
    z = pseudorandom(700, 300, seed=3)
    ref = oracle.kramers_kronig_matrix(np.asarray(g)) @ z
    outs = []
    for shape in range(5):
        os.environ["EPSX_GEMM_SHAPE"] = str(shape)
        try:
            outs.append(epsx.kramers_kronig(g, z, ctx))
        finally:
            del os.environ["EPSX_GEMM_SHAPE"]
    assert rel(outs[0], ref) <= 1e-13
    for k, y in enumerate(outs[1:], 1):
        assert np.array_equal(y, outs[0]), k
    # k = 200 (7 k-tiles of 32): no shape splits k, so the first eight shapes (the last of them
    # the register-pipelined 256x32 kernel with one wave along k) give the same bits; 300 columns
    # leave a partial tile
    g = FrequencyGrid(-1.0, 0.01, np.arange(200))
    z = pseudorandom(200, 300, seed=4)
    ref = oracle.kramers_kronig_matrix(np.asarray(g)) @ z
    outs = []
    for shape in range(12):
        os.environ["EPSX_GEMM_SHAPE"] = str(shape)
        try:
            outs.append(epsx.kramers_kronig(g, z, ctx))
        finally:
            del os.environ["EPSX_GEMM_SHAPE"]
    assert rel(outs[0], ref) <= 1e-13
    for k, y in enumerate(outs[1:8], 1):
        assert np.array_equal(y, outs[0]), k
    # shape 8 (two waves along k: the even and the odd k-tiles summed apart, then added) and the
    # 16-byte-load shapes 9, 10 and 11 (rows and k paired; 2 or 4 waves along k, or 2 x 2 waves over
    # a 128x128 tile; the pitches here are even) have fixed orders of their own: the oracle at 1e-13,
    # the same bits on a repeat, and on the 700-node grid (k splits, several row tiles)
    g7 = FrequencyGrid(-1.0, 0.003, np.arange(700))
    z7 = pseudorandom(700, 300, seed=3)
    ref7 = oracle.kramers_kronig_matrix(np.asarray(g7)) @ z7
    for shape in (8, 9, 10, 11):
        assert rel(outs[shape], ref) <= 1e-13, shape
        os.environ["EPSX_GEMM_SHAPE"] = str(shape)
        try:
            assert np.array_equal(epsx.kramers_kronig(g, z, ctx), outs[shape]), shape
            assert rel(epsx.kramers_kronig(g7, z7, ctx), ref7) <= 1e-13, shape
        finally:
            del os.environ["EPSX_GEMM_SHAPE"]
    # an odd pitch (77 nodes): the 16-byte-load shapes fall back to the one-wave register kernel
    g3 = FrequencyGrid(0.0, 0.01, np.arange(77))
    z3 = pseudorandom(77, 3, seed=2)
    ref3 = oracle.kramers_kronig_matrix(np.asarray(g3)) @ z3
    for shape in (9, 10, 11):
        os.environ["EPSX_GEMM_SHAPE"] = str(shape)
        try:
            assert rel(epsx.kramers_kronig(g3, z3, ctx), ref3) <= 1e-13, shape
        finally:
            del os.environ["EPSX_GEMM_SHAPE"]


def kernel_case():
    """The data of test/test_scgd0_kernel.jl:6-31."""
    nm, ni, nf = 3, 13, 17
    ωs = FrequencyGrid.from_nodes(np.concatenate([np.linspace(-1.0, -0.6, 21), np.linspace(-0.5, 0.5, 11),
                                                  np.linspace(0.6, 1.0, 21)]))
    nω = len(ωs)
    Sin = 0.1 * pseudorandom(nω, ni, seed=7) + 1j * (-0.05 - 0.05 * np.abs(pseudorandom(nω, ni, seed=8)))
    g2 = np.abs(pseudorandom(nm, ni, nf, seed=9))
    wq = 0.05 + 0.2 * np.abs(pseudorandom(nm, ni, nf, seed=10))
    w_ac = 1e-3
    wq[1, 2, :] = 0.0
    g2[1, 2, :] = np.inf
    e_f = 0.8 * pseudorandom(nf, seed=11)
    f_to_i = np.array([(5 * f) % ni + 1 for f in range(1, nf + 1)], dtype=np.int32)
    w_f = np.full(nf, 1.0 / nf)
    return dict(ωs=ωs, Sin=Sin, g2=g2, wq=wq, w_ac=w_ac, e_f=e_f, f_to_i=f_to_i, w_f=w_f, mu=0.07, T=0.03)


def run_fm_device(ctx, c, T=None, precision=0):
    from epsx import _lib
    nm, ni, nf = c["g2"].shape
    ws = c["ωs"]
    ImS = np.full((len(ws), ni), 0.5, order="F")
    g2 = np.asfortranarray(c["g2"])
    wq = np.asfortranarray(c["wq"])
    off = np.ascontiguousarray(ws.offsets)
    Sin = np.asfortranarray(c["Sin"], dtype=np.complex128)
    e_f = np.ascontiguousarray(c["e_f"], dtype=np.float64)
    w_f = np.ascontiguousarray(c["w_f"], dtype=np.float64)
    f_to_i = np.ascontiguousarray(c["f_to_i"], dtype=np.int32)
    ctx.check(_lib.lib().epsx_add_fan_migdal_imsigma(
        ctx.handle, nm, ni, nf, _lib.ptr(g2), _lib.ptr(wq), _lib.ptr(e_f),
        _lib.ptr(f_to_i), _lib.ptr(w_f), len(ws), ws.ω_origin, ws.dω_unit,
        _lib.ptr(off), Sin.shape[1], _lib.ptr(Sin), c["mu"], c["T"] if T is None else T, c["w_ac"],
        precision, _lib.ptr(ImS)))
    return ImS


def run_fm_oracle(c, T=None):
    nm, ni, nf = c["g2"].shape
    ws = np.asarray(c["ωs"])
    table = oracle.cell_search_table(ws, c["ωs"].dω_min)
    wq_fi = np.asfortranarray(np.transpose(c["wq"], (0, 2, 1)))
    ImS = np.full((len(ws), ni), 0.5, order="F")
    return oracle.add_fan_migdal(ImS, c["g2"], wq_fi, c["e_f"], c["f_to_i"], c["w_f"], ws, table, c["Sin"],
                                 c["mu"], c["T"] if T is None else T, c["w_ac"])


def test_fan_migdal_kernel_fp64(ctx):
    c = kernel_case()
    ImS = run_fm_device(ctx, c)
    ref = run_fm_oracle(c)
    naive = naive_fan_migdal(np.full_like(ref, 0.5), np.asarray(c["ωs"]), c["Sin"], c["g2"], c["wq"], c["e_f"],
                             c["f_to_i"], c["w_f"], c["mu"], c["T"], c["w_ac"])
    assert np.all(np.isfinite(ImS))
    assert rel(ImS, ref) <= 1e-13
    assert rel(ImS, naive) <= 1e-13
    # run to run: bitwise
    assert np.array_equal(run_fm_device(ctx, c), ImS)


def test_fan_migdal_kernel_step_temperature(ctx):
    # T = 0: the step occupations of src/utils.jl:18-21 and :29-36
    c = kernel_case()
    ImS = run_fm_device(ctx, c, T=0.0)
    ref = run_fm_oracle(c, T=0.0)
    assert np.all(np.isfinite(ImS))
    assert rel(ImS, ref) <= 1e-13


def test_fan_migdal_kernel_fp32(ctx):
    c = kernel_case()
    ImS = run_fm_device(ctx, c, precision=1)
    ref = run_fm_oracle(c)
    # FP32 terms with a compensated sum over f; stated tolerance 1e-5 relative
    assert rel(ImS - 0.5, ref - 0.5) <= 1e-5


def holstein(nk=64, nodes=None, T=(0.1,), nlist=0.0, μ_start=None):
    ωs = FrequencyGrid.from_nodes(np.arange(-200, 201) * 0.02 if nodes is None else nodes)
    if μ_start is None:
        μ_start = [holstein_bare_mu(nk=nk, t=1.0, T=Ti, n_target=1 + nlist, spin_degeneracy=2) for Ti in T]
    return holstein_1d_setup(nk, 1.0, 0.2, 0.2, ωs, μ_start)


def occ(T, nlist=0.0, typ="Metal", μlist=None):
    return epsx.Occupation(Tlist=np.asarray(T, dtype=float), nlist=nlist, nelec=1.0, spin_degeneracy=2.0,
                           type=typ, μlist=μlist)


def run_both(s, o, **kw):
    kw.setdefault("η_min", 0.0)
    kw.setdefault("ω_acoustic", 1e-6)
    dev_kw = {k: v for k, v in kw.items() if k not in ("semiconductor",)}
    r = epsx.loop_scgd0(s, o, **dev_kw)
    okw = {k: v for k, v in kw.items() if k not in ("method", "precision", "batched")}
    ref = oracle.loop_scgd0(s, o.Tlist, o.nlist, o.nelec, semiconductor=(o.type == "Semiconductor"),
                            spin_degeneracy=o.spin_degeneracy, **okw)
    return r, ref


@pytest.mark.parametrize("nlist", [0.0, -0.3])
def test_loop_holstein_self_consistent(ctx, nlist):
    s = holstein(nlist=nlist)
    r, ref = run_both(s, occ([0.1], nlist), η_init=0.05, maxiter=30, tol=1e-6)
    assert r["converged"].all() and ref["converged"].all()
    assert r["history"]["err"].shape == ref["history"]["err"].shape
    assert rel(r["Σ"], ref["Σ"]) <= 1e-10
    assert abs(r["μ"][0] - ref["μ"][0]) <= 1e-10 * max(1.0, abs(ref["μ"][0]))
    assert np.allclose(r["history"]["min_spectral_weight"], ref["history"]["min_spectral_weight"], rtol=1e-10)
    # k-independence of the Holstein Σ (test/test_scgd0_holstein.jl:39)
    assert np.max(np.abs(r["Σ"] - r["Σ"][:, :1, :])) < 1e-12


def test_loop_two_temperatures_and_batching(ctx):
    s = holstein(T=(0.1, 0.2))
    o = occ([0.1, 0.2])
    r, ref = run_both(s, o, η_init=0.05, maxiter=30, tol=1e-6)
    assert rel(r["Σ"], ref["Σ"]) <= 1e-10
    assert np.array_equal(r["history"]["iter_frozen"], ref["history"]["iter_frozen"])
    rb = epsx.loop_scgd0(s, o, η_min=0.0, η_init=0.05, maxiter=30, tol=1e-6, batched=True)
    assert rel(rb["Σ"], r["Σ"]) <= 1e-12
    assert np.array_equal(rb["history"]["iter_frozen"], r["history"]["iter_frozen"])
    # each temperature alone (test/test_scgd0_holstein.jl:77-85)
    for iT, T in enumerate((0.1, 0.2)):
        s1 = holstein(T=(T,))
        r1 = epsx.loop_scgd0(s1, occ([T]), η_min=0.0, η_init=0.05, maxiter=30, tol=1e-6)
        assert rel(r["Σ"][:, :, iT], r1["Σ"][:, :, 0]) <= 1e-12
        assert abs(r["μ"][iT] - r1["μ"][0]) < 1e-12


def test_loop_one_shot_fixed_mu_and_tiling(ctx):
    s = holstein(μ_start=[0.3], nlist=-0.3)
    o = occ([0.1], -0.3, μlist=[0.3])
    r, ref = run_both(s, o, η_init=0.05, maxiter=1, tol=1e-6, fix_μ=True)
    assert rel(r["Σ"], ref["Σ"]) <= 1e-12
    assert r["μ"][0] == 0.3 and not r["converged"][0]
    rt = epsx.loop_scgd0(s, o, η_min=0.0, η_init=0.05, maxiter=1, tol=1e-6, fix_μ=True, gpu_tile=7)
    assert rel(rt["Σ"], r["Σ"]) <= 1e-12


def test_loop_restart_converges_in_one(ctx):
    s = holstein()
    o = occ([0.1])
    r = epsx.loop_scgd0(s, o, η_min=0.0, η_init=0.05, maxiter=30, tol=1e-6)
    r2 = epsx.loop_scgd0(s, o, η_min=0.0, η_init=0.05, maxiter=30, tol=1e-6, Σ_init=r["Σ"], μ_start=r["μ"])
    assert r2["history"]["err"].shape[1] == 1 and r2["converged"].all()
    assert np.max(np.abs(r2["Σ"] - r["Σ"])) < 1e-6
    assert abs(r2["μ"][0] - r["μ"][0]) < 1e-6


def test_loop_maxiter_no_root_and_multiplets(ctx):
    s = holstein()
    with pytest.warns(UserWarning, match="not converged"):
        r = epsx.loop_scgd0(s, occ([0.1]), η_min=0.0, η_init=0.05, maxiter=2, tol=1e-6)
    assert not r["converged"][0] and r["history"]["iter_frozen"][0] == 0 and r["history"]["err"].shape[1] == 2
    with pytest.raises(epsx.ArgumentError, match="no chemical potential"):
        epsx.loop_scgd0(s, occ([0.1], 1.5), η_min=0.0, η_init=0.05, maxiter=20, tol=1e-6)
    # multiplet average (test/test_scgd0_loop.jl:25-37)
    import dataclasses
    nk = 64
    g2v = np.array([[0.2 * (1 + 0.1 * i) for _ in range(nk)] for i in range(1, nk // 2 + 2)]).reshape(1, nk // 2 + 1, nk)
    sv = dataclasses.replace(s, g2=np.asfortranarray(g2v))
    kw = dict(η_min=0.0, η_init=0.05, maxiter=1, tol=1e-6, fix_μ=True)
    o = occ([0.1], μlist=[0.0])
    ra = epsx.loop_scgd0(sv, o, **kw)
    rb = epsx.loop_scgd0(dataclasses.replace(sv, multiplets=[[2, 3]]), o, **kw)
    assert np.linalg.norm(ra["Σ"][:, 1, 0] - ra["Σ"][:, 2, 0]) > 1e-3 * np.linalg.norm(ra["Σ"][:, 1, 0])
    assert np.array_equal(rb["Σ"][:, 1, 0], rb["Σ"][:, 2, 0])
    assert rel(rb["Σ"][:, 1, 0], (ra["Σ"][:, 1, 0] + ra["Σ"][:, 2, 0]) / 2) <= 1e-14
    keep = [0] + list(range(3, nk // 2 + 1))
    assert np.array_equal(rb["Σ"][:, keep, :], ra["Σ"][:, keep, :])
    ref = oracle.loop_scgd0(dataclasses.replace(sv, multiplets=[[2, 3]]), [0.1], [0.0], 1.0, η_min=0.0, η_init=0.05,
                            maxiter=1, tol=1e-6, ω_acoustic=1e-6, fix_μ=True, μ_start=[0.0])
    assert rel(rb["Σ"], ref["Σ"]) <= 1e-12


def test_loop_run_to_run_bitwise(ctx):
    s = holstein(T=(0.1, 0.2))
    o = occ([0.1, 0.2])
    a = epsx.loop_scgd0(s, o, η_min=0.0, η_init=0.05, maxiter=4, tol=0.0)
    b = epsx.loop_scgd0(s, o, η_min=0.0, η_init=0.05, maxiter=4, tol=0.0)
    assert np.array_equal(a["Σ"], b["Σ"]) and np.array_equal(a["μ"], b["μ"])


def test_loop_holstein_c1(ctx):
    """C1: nk = 128, ωs = -4:0.005:4 (1601 nodes), η_init = 0.01, T = 0.1 and 0.3, tol 1e-8."""
    T = (0.1, 0.3)
    s = holstein(nk=128, nodes=np.arange(-800, 801) * 0.005, T=T)
    r, ref = run_both(s, occ(list(T)), η_init=0.01, maxiter=100, tol=1e-8)
    assert r["converged"].all() and ref["converged"].all()
    assert np.array_equal(r["history"]["iter_frozen"], ref["history"]["iter_frozen"])
    assert rel(r["Σ"], ref["Σ"]) <= 1e-10
    assert np.all(np.abs(r["μ"] - ref["μ"]) <= 1e-10 * np.maximum(1.0, np.abs(ref["μ"])))


def test_loop_t2g_small_vs_oracle(ctx):
    """The synthetic t2g model at nk = 8 (multiplets, 9 Einstein modes, graded grid) against the
    oracle for 3 iterations with the μ update."""
    s, meta = t2g_setup(nk=8, Tlist_K=(100.0, 300.0), grid_windows=((0.3, 4e-3), (1.2, 16e-3), (3.0, 64e-3)))
    assert len(s.multiplets) > 0
    o = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r, ref = run_both(s, o, η_init=5e-2, maxiter=3, tol=0.0, ω_acoustic=1e-6)
    assert rel(r["Σ"], ref["Σ"]) <= 1e-10
    assert np.all(np.abs(r["μ"] - ref["μ"]) <= 1e-10 * np.abs(ref["μ"]))
    for m in s.multiplets:
        for k in m[1:]:
            assert np.array_equal(r["Σ"][:, k - 1, :], r["Σ"][:, m[0] - 1, :])


def test_star_product_blaslt_matches_own_kernel(ctx):
    """The star method's FP64 products in hipBLASLt (EPSX_STAR_BLAS=1) against the library's own
    MFMA kernel (the default): batched runs of the Holstein chain at two temperatures that freeze at
    different iterations (the replayed graph captured again for one temperature), same iterations
    and Σ to 1e-12; and the t2g model at nk = 8 (multiplets: the product over the class columns,
    stored at an even pitch) for three iterations, Σ to 1e-12."""
    s = holstein(T=(0.1, 0.2))
    o = occ([0.1, 0.2])
    kw = dict(η_min=0.0, η_init=0.05, maxiter=30, tol=1e-6, batched=True, method="star")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        own = epsx.loop_scgd0(s, o, **kw)
        os.environ["EPSX_STAR_BLAS"] = "1"
        try:
            lt = epsx.loop_scgd0(s, o, **kw)
        finally:
            del os.environ["EPSX_STAR_BLAS"]
        # EPSX_STAR_BLAS=0: no library product at all (Kramers-Kronig in the GEMM's complex epilogue)
        os.environ["EPSX_STAR_BLAS"] = "0"
        try:
            ep = epsx.loop_scgd0(s, o, **kw)
        finally:
            del os.environ["EPSX_STAR_BLAS"]
    assert lt["converged"].all() and own["converged"].all() and ep["converged"].all()
    fr = lt["history"]["iter_frozen"]
    assert fr[0] != fr[1] and np.array_equal(fr, own["history"]["iter_frozen"])
    assert np.array_equal(fr, ep["history"]["iter_frozen"])
    for r in (lt, ep):
        assert rel(r["Σ"], own["Σ"]) <= 1e-12
        assert np.all(np.abs(r["μ"] - own["μ"]) <= 1e-12)
    s2, meta = t2g_setup(nk=8, Tlist_K=(100.0, 300.0), grid_windows=((0.3, 4e-3), (1.2, 16e-3), (3.0, 64e-3)))
    o2 = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    kw2 = dict(η_min=0.0, η_init=5e-2, maxiter=3, tol=0.0, ω_acoustic=1e-6, batched=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dev = epsx.Scgd0Device(s2, ctx)
        own2 = dev.loop(o2, method="star", **kw2)
        os.environ["EPSX_STAR_BLAS"] = "1"
        try:
            lt2 = dev.loop(o2, method="star", **kw2)
        finally:
            del os.environ["EPSX_STAR_BLAS"]
    assert rel(lt2["Σ"], own2["Σ"]) <= 1e-12
    assert np.all(np.abs(lt2["μ"] - own2["μ"]) <= 1e-12)


def test_star_product_blaslt_batch_split(ctx):
    """The hipBLASLt products (EPSX_STAR_BLAS=1) over five temperatures split into chunks of at most
    two instances (EPSX_LT_CHUNK=2: calls of 2 + 2 + 1, each chunk's pointers advanced by its offset)
    against the library's own MFMA kernel: same iterations, Σ and μ to 1e-12."""
    T = (0.08, 0.1, 0.13, 0.16, 0.2)
    s = holstein(T=T)
    o = occ(list(T))
    kw = dict(η_min=0.0, η_init=0.05, maxiter=80, tol=1e-6, batched=True, method="star")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        os.environ["EPSX_LT_CHUNK"] = "2"
        os.environ["EPSX_STAR_BLAS"] = "1"
        try:
            lt = epsx.loop_scgd0(s, o, **kw)
        finally:
            del os.environ["EPSX_LT_CHUNK"]
            del os.environ["EPSX_STAR_BLAS"]
        own = epsx.loop_scgd0(s, o, **kw)
    assert lt["converged"].all() and own["converged"].all()
    assert np.array_equal(lt["history"]["iter_frozen"], own["history"]["iter_frozen"])
    assert rel(lt["Σ"], own["Σ"]) <= 1e-12
    assert np.all(np.abs(lt["μ"] - own["μ"]) <= 1e-12)


def test_star_method_matches_direct_and_oracle(ctx):
    """The factorised Fan-Migdal method (Einstein modes, ε_f == ε_{i_f}) against the direct kernel
    and the oracle on the t2g model at nk = 8 (multiplets, 9 modes, graded grid)."""
    s, meta = t2g_setup(nk=8, Tlist_K=(100.0, 300.0), grid_windows=((0.3, 4e-3), (1.2, 16e-3), (3.0, 64e-3)))
    o = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    kw = dict(η_min=0.0, η_init=5e-2, maxiter=3, tol=0.0, ω_acoustic=1e-6)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dev = epsx.Scgd0Device(s, ctx)
        rd = dev.loop(o, method="direct", **kw)
        rs = dev.loop(o, method="star", **kw)
        ra = dev.loop(o, method="auto", batched=True, **kw)
        ref = oracle.loop_scgd0(s, o.Tlist, o.nlist, 1.0, **kw)
    assert rs["info"]["fm_method"] == 2 and ra["info"]["fm_method"] == 2 and rd["info"]["fm_method"] == 0
    # the coupling reduction kept from the previous run gives the same bits
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        rr = dev.loop(o, method="star", reuse_star=True, **kw)
    assert np.array_equal(rr["Σ"], rs["Σ"]) and rr["info"]["timings"]["star_reduce_launches"] == 0
    assert rel(rs["Σ"], rd["Σ"]) <= 1e-12
    assert rel(rs["Σ"], ref["Σ"]) <= 1e-10
    assert np.all(np.abs(rs["μ"] - ref["μ"]) <= 1e-10 * np.abs(ref["μ"]))
    assert np.array_equal(ra["Σ"], rs["Σ"])
    # every star read from global memory (the path of stars too large for the LDS stage)
    os.environ["EPSX_STAR_STAGE"] = "global"
    try:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            rg = epsx.Scgd0Device(s, ctx).loop(o, method="star", **kw)
    finally:
        del os.environ["EPSX_STAR_STAGE"]
    assert rel(rg["Σ"], rs["Σ"]) <= 1e-14
    # the precondition is checked: a dispersive mode refuses the star method
    import dataclasses
    s2 = dataclasses.replace(s, ωph=np.asfortranarray(np.hstack([s.ωph, s.ωph * 1.01])),
                             iq_kk=np.asfortranarray(np.array([[1, 2]], dtype=np.int32)),
                             iks_f=np.asfortranarray(1 + (np.arange(len(s.ε_f)) % 2).astype(np.int32)))
    with pytest.raises(epsx.ArgumentError, match="star"):
        epsx.loop_scgd0(s2, o, method="star", ctx=ctx, **kw)
    r2 = epsx.loop_scgd0(s2, o, method="auto", ctx=ctx, **kw)
    assert r2["info"]["fm_method"] == 0


def test_star_method_holstein_fine_grid(ctx):
    """The star method on the Holstein chain (one Einstein mode ω_ph = 0.2, inner states k and
    their representatives min(j, N - j)) on a 2 meV grid, where ω ± ω_ph lands 100 cells away:
    against the direct kernel, 1e-12."""
    s = holstein(nk=32, nodes=np.arange(-1250, 1251) * 0.002)
    o = occ([0.1])
    kw = dict(η_min=0.0, η_init=0.05, maxiter=3, tol=0.0, ω_acoustic=1e-6, fix_μ=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dev = epsx.Scgd0Device(s, ctx)
        rs = dev.loop(o, method="star", **kw)
        rd = dev.loop(o, method="direct", **kw)
    assert rs["info"]["fm_method"] == 2
    assert rel(rs["Σ"], rd["Σ"]) <= 1e-12


@pytest.mark.parametrize("prec,method", [("fp32", "star"), ("bf16", "star"), ("fp32", "direct")])
def test_low_precision_phase_then_fp64_sweep(ctx, prec, method):
    """FP32 / BF16 products (star) or FP32 terms (direct) until the low-precision floor, then FP64
    iterations to the same tol: the result agrees with the all-FP64 run to the tolerance scale
    (both are within tol of the fixed point), in the batched graph loop and the one-by-one loop."""
    s, meta = t2g_setup(nk=8, Tlist_K=(100.0, 300.0), grid_windows=((0.3, 4e-3), (1.2, 16e-3), (3.0, 64e-3)))
    o = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    kw = dict(η_min=0.0, η_init=5e-2, maxiter=200, tol=1e-7, ω_acoustic=1e-6, method=method)
    dev = epsx.Scgd0Device(s, ctx)
    ref = dev.loop(o, batched=True, **kw)
    assert ref["converged"].all() and ref["info"]["n_low_precision"] == 0
    for batched in (True, False):
        r = dev.loop(o, batched=batched, precision=prec, **kw)
        assert r["converged"].all()
        nlp, n = r["info"]["n_low_precision"], r["history"]["err"].shape[1]
        assert 1 <= nlp < n
        # the last low-precision iterate is farther from the fixed point than tol allows, the FP64
        # sweep brings it back: the final residual is below tol
        assert np.all(r["history"]["err"][np.arange(len(o.Tlist)), r["history"]["iter_frozen"] - 1] < kw["tol"])
        scale = np.max(np.abs(ref["Σ"]))
        assert np.max(np.abs(r["Σ"] - ref["Σ"])) <= 50 * kw["tol"] + 1e-9 * scale
        assert np.max(np.abs(r["μ"] - ref["μ"])) <= 1e-6


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_low_precision_star_product_native_matches_staged(ctx, prec):
    """The BF16 / FP32 star product and Kramers-Kronig off-diagonal from operands stored k-contiguous
    in that precision (h built in it, G̃ and KK converted once, ImΣ per iteration) against the
    products that round the FP64 operands as they stage them: the same rounded operands, so the two
    agree to FP32 accumulation order (1e-5 relative), while both differ from FP64 by the operand
    rounding.  One iteration, in the low-precision phase."""
    s, meta = t2g_setup(nk=8, Tlist_K=(100.0, 300.0), grid_windows=((0.3, 4e-3), (1.2, 16e-3), (3.0, 64e-3)))
    o = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    kw = dict(η_min=0.0, η_init=5e-2, maxiter=1, tol=0.0, ω_acoustic=1e-6, method="star", fix_μ=True)
    dev = epsx.Scgd0Device(s, ctx)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r64 = dev.loop(o, batched=True, **kw)
        rn = dev.loop(o, batched=True, precision=prec, **kw)
        os.environ["EPSX_LP_STAGED"] = "1"
        try:
            rs = dev.loop(o, batched=True, precision=prec, **kw)
        finally:
            del os.environ["EPSX_LP_STAGED"]
    assert rn["info"]["n_low_precision"] == 1 and rs["info"]["n_low_precision"] == 1
    d_ns = rel(rn["Σ"].imag, rs["Σ"].imag)
    d_64 = rel(rn["Σ"].imag, r64["Σ"].imag)
    assert d_ns <= (1e-5 if prec == "bf16" else 1e-6), d_ns
    # Re Σ: Kramers-Kronig's off-diagonal from the stored low-precision operands as well
    d_re = rel(rn["Σ"].real, rs["Σ"].real)
    assert d_re <= (1e-5 if prec == "bf16" else 1e-6), d_re
    lo, hi = (1e-7, 1e-2) if prec == "bf16" else (1e-12, 1e-5)
    assert lo < d_64 < hi, d_64


V2_T, V2_ETA = [0.03, 0.06], [0.0, 0.03]


@pytest.mark.parametrize("identity", [True, False])
def test_loop_dispersive_direct_eta_min_vs_oracle(ctx, identity):
    """The direct Fan-Migdal method on a dispersive toy: ωph[ν, q] varying with q (one entry below
    ω_acoustic), a non-square, non-symmetric iq_kk and iks_i != iks_f, so the device's per-pair
    gather ωph[ν, iq_kk[iks_i[i], iks_f[f]]] is what the result depends on; per-temperature η_min =
    [0, 0.03] (test/pb_fixture.jl:16); three iterations with the μ update (test/test_scgd0_gpu.jl:
    7-37).  Against the oracle at 1e-10 (Σ and μ), one by one and batched, and, for f -> i the
    identity, against the v2 iteration over the plain-loop ωq (test/test_scgd0_pb.jl:70-104)."""
    from toys import dispersive_setup, pair_frequencies
    from ref_numpy import scgd0_v2_reference
    s = dispersive_setup(identity=identity)
    o = occ(V2_T)
    mu0 = [s.μ_start[0]] * 2
    kw = dict(η_min=V2_ETA, η_init=0.05, maxiter=3, tol=0.0, ω_acoustic=1e-6, μ_start=mu0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dev = epsx.Scgd0Device(s, ctx)
        r = dev.loop(o, method="auto", **kw)
        rb = dev.loop(o, method="direct", batched=True, **kw)
        ra = dev.loop(o, method="direct", **kw)
    ref = oracle.loop_scgd0(s, V2_T, 0.0, 1.0, **kw)
    assert r["info"]["fm_method"] == 0 and rb["info"]["fm_method"] == 0
    for x in (r, rb):
        assert rel(x["Σ"], ref["Σ"]) <= 1e-10
        assert np.all(np.abs(x["μ"] - ref["μ"]) <= 1e-10 * np.abs(ref["μ"]))
    assert np.array_equal(ra["Σ"], r["Σ"])
    if identity:
        ws = np.asarray(s.ωs)
        Σ2, μ2 = scgd0_v2_reference(s.ε_i, s.w_i, s.nstates_base, s.g2, pair_frequencies(s), ws,
                                    np.asarray(s.ωs_spectral), oracle.kramers_kronig_matrix(ws), V2_T, [0.0, 0.0],
                                    1.0, 2.0, mu0, V2_ETA, 0.05, 3, 1e-6)
        assert rel(r["Σ"], Σ2) <= 1e-10
        assert np.all(np.abs(r["μ"] - μ2) <= 1e-9)


@pytest.mark.parametrize("nlist", [-0.3, 0.3])
def test_loop_semiconductor_mu_branches(ctx, nlist):
    """occ.type = :Semiconductor: the hole count for nlist < 0 and the electron count for
    nlist > 0 (src/scgd0.jl:314, :361-366) in the device μ solve, three iterations, against the
    oracle at 1e-10; and the hole branch gives another μ than the electron formula would."""
    s = holstein(nlist=nlist)
    o = occ([0.1], nlist, typ="Semiconductor")
    r, ref = run_both(s, o, η_init=0.05, maxiter=3, tol=0.0)
    assert rel(r["Σ"], ref["Σ"]) <= 1e-10
    assert abs(r["μ"][0] - ref["μ"][0]) <= 1e-10 * max(1.0, abs(ref["μ"][0]))
    rm, _ = run_both(s, occ([0.1], nlist, typ="Metal"), η_init=0.05, maxiter=3, tol=0.0)
    if nlist < 0:
        # the window's Σ w_i ∫A_i is not exactly 1, so holes and electrons give different μ
        assert abs(rm["μ"][0] - r["μ"][0]) > 1e-9
    else:
        assert rel(rm["Σ"], r["Σ"]) <= 1e-13 and rm["μ"][0] == r["μ"][0]


def test_star_method_eta_min_clamp(ctx):
    """The star method's h build clamps the representative's Σ at -η_min[T]: the t2g model at
    nk = 8 with η_min = [0, 0.03] eV, star against direct (1e-12) and the oracle (1e-10)."""
    s, meta = t2g_setup(nk=8, Tlist_K=(100.0, 300.0), grid_windows=((0.3, 4e-3), (1.2, 16e-3), (3.0, 64e-3)))
    o = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    kw = dict(η_min=[0.0, 0.03], η_init=5e-2, maxiter=3, tol=0.0, ω_acoustic=1e-6)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dev = epsx.Scgd0Device(s, ctx)
        rs = dev.loop(o, method="star", batched=True, **kw)
        rd = dev.loop(o, method="direct", **kw)
        r0 = dev.loop(o, method="star", batched=True, **dict(kw, η_min=0.0))
    ref = oracle.loop_scgd0(s, o.Tlist, o.nlist, 1.0, **kw)
    assert rs["info"]["fm_method"] == 2
    assert rel(rs["Σ"], rd["Σ"]) <= 1e-12
    assert rel(rs["Σ"], ref["Σ"]) <= 1e-10
    assert np.all(np.abs(rs["μ"] - ref["μ"]) <= 1e-10 * np.abs(ref["μ"]))
    assert np.array_equal(rs["Σ"][:, :, 0], r0["Σ"][:, :, 0]) or rel(rs["Σ"][:, :, 0], r0["Σ"][:, :, 0]) <= 1e-13
    assert rel(rs["Σ"][:, :, 1], r0["Σ"][:, :, 1]) > 1e-4


def test_c3_t2g24_eight_temperatures_mixed_precision(ctx):
    """C3: the 24³ t2g model (g2 1.85 GB resident), 8 temperatures 50..400 K in one batched,
    graph-replayed run with per-temperature masks, FP64 and the BF16 / FP32 level schedules (a final
    FP64 sweep to the same tol).  Every temperature converges to tol = 1e-6 eV; the batched FP64 run
    freezes each temperature at the iteration of the one-by-one loop, with the same Σ to 1e-12; the
    mixed-precision Σ is within 1e-5 relative (and μ within 1e-6 eV) of the FP64 solve; multiplet
    members are bitwise equal."""
    temps = tuple(50.0 * k for k in range(1, 9))
    s, meta = t2g_setup(nk=24, Tlist_K=temps)
    o = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    kw = dict(η_min=0.0, η_init=5e-3, maxiter=200, tol=1e-6, ω_acoustic=1e-6, method="auto")
    dev = epsx.Scgd0Device(s, ctx)
    r64 = dev.loop(o, batched=True, **kw)
    assert r64["converged"].all() and r64["info"]["fm_method"] == 2
    fr = r64["history"]["iter_frozen"]
    assert len(set(fr.tolist())) > 1          # the temperatures stop at different iterations
    one = dev.loop(o, batched=False, **kw)
    assert np.array_equal(one["history"]["iter_frozen"], fr)
    assert rel(one["Σ"], r64["Σ"]) <= 1e-12
    scale = np.max(np.abs(r64["Σ"]))
    for prec in ("bf16", "fp32"):
        r = dev.loop(o, batched=True, precision=prec, **kw)
        assert r["converged"].all() and r["info"]["n_low_precision"] >= 1, prec
        assert np.max(np.abs(r["Σ"] - r64["Σ"])) <= 1e-5 * scale, prec
        assert np.max(np.abs(r["μ"] - r64["μ"])) <= 1e-6, prec
        for m in s.multiplets:
            for k in m[1:]:
                assert np.array_equal(r["Σ"][:, k - 1, :], r["Σ"][:, m[0] - 1, :])


def einstein_toy():
    """The dispersive toy with f -> i not the identity and ε_f != ε_{i_f}, its modes made
    dispersionless (ωph[ν, :] = ωph[ν, 4]): the Einstein factorisation applies, the star method does
    not."""
    import dataclasses
    from toys import dispersive_setup
    s = dispersive_setup(identity=False)
    nm, nq = s.ωph.shape
    return dataclasses.replace(s, ωph=np.asfortranarray(np.repeat(s.ωph[:, 4:5], nq, axis=1)))


def test_einstein_method_matches_direct_and_oracle(ctx):
    """The Einstein factorisation ImΣ = H' g2 (dispersionless modes, inner-state energies that
    differ from their representatives') against the direct kernel (1e-12) and the oracle (1e-10),
    with per-temperature η_min and the μ update: picked by the automatic choice, the same one by one,
    batched and in chunks of 3 inner states (gpu_tile, which only reorders the sum over f)."""
    s = einstein_toy()
    o = occ(V2_T)
    kw = dict(η_min=V2_ETA, η_init=0.05, maxiter=3, tol=0.0, ω_acoustic=1e-6, μ_start=[s.μ_start[0]] * 2)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dev = epsx.Scgd0Device(s, ctx)
        ra = dev.loop(o, method="auto", **kw)
        rd = dev.loop(o, method="direct", **kw)
        rb = dev.loop(o, method="einstein", batched=True, **kw)
        rt = dev.loop(o, method="einstein", gpu_tile=3, **kw)
        with pytest.raises(epsx.ArgumentError, match="star"):
            dev.loop(o, method="star", **kw)
    ref = oracle.loop_scgd0(s, V2_T, 0.0, 1.0, **kw)
    assert ra["info"]["fm_method"] == 1 and rb["info"]["fm_method"] == 1 and rd["info"]["fm_method"] == 0
    assert rel(ra["Σ"], rd["Σ"]) <= 1e-12
    assert rel(ra["Σ"], ref["Σ"]) <= 1e-10
    assert np.all(np.abs(ra["μ"] - ref["μ"]) <= 1e-10 * np.abs(ref["μ"]))
    assert rel(rb["Σ"], ra["Σ"]) <= 1e-13 and rel(rt["Σ"], ra["Σ"]) <= 1e-12


def test_einstein_method_acoustic_modes(ctx):
    """A mode below ω_acoustic is a zero column of H': with finite couplings there the Einstein method
    runs and equals the direct kernel; with Inf couplings there (allowed: the term-by-term sum skips
    the mode before reading g2, src/scgd0_kernel.jl:22-25) the automatic choice falls back to the
    direct kernel and an explicit request is refused."""
    import dataclasses
    s = einstein_toy()
    wph = s.ωph.copy()
    wph[0, :] = 0.0
    s0 = dataclasses.replace(s, ωph=np.asfortranarray(wph))
    o = occ(V2_T)
    kw = dict(η_min=V2_ETA, η_init=0.05, maxiter=2, tol=0.0, ω_acoustic=1e-6, μ_start=[s.μ_start[0]] * 2,
              fix_μ=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dev = epsx.Scgd0Device(s0, ctx)
        re_ = dev.loop(o, method="einstein", **kw)
        rd = dev.loop(o, method="direct", **kw)
        assert re_["info"]["fm_method"] == 1 and rel(re_["Σ"], rd["Σ"]) <= 1e-12
        g2 = s0.g2.copy()
        g2[0] = np.inf
        dev = epsx.Scgd0Device(dataclasses.replace(s0, g2=np.asfortranarray(g2)), ctx)
        ra = dev.loop(o, method="auto", **kw)
        assert ra["info"]["fm_method"] == 0 and np.all(np.isfinite(ra["Σ"]))
        assert rel(ra["Σ"], rd["Σ"]) <= 1e-12
        with pytest.raises(epsx.ArgumentError, match="finite couplings"):
            dev.loop(o, method="einstein", **kw)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_einstein_low_precision_then_fp64(ctx, prec):
    """The t2g model at nk = 8 with the inner-state energies moved off their representatives' by up
    to 1e-6 eV (the star method's precondition fails, the Einstein one holds): the automatic choice
    runs the Einstein factorisation; its FP32 / BF16 products until the low-precision floor, then FP64
    to the same tol, reach the FP64 fixed point."""
    import dataclasses
    s, meta = t2g_setup(nk=8, Tlist_K=(100.0, 300.0), grid_windows=((0.3, 4e-3), (1.2, 16e-3), (3.0, 64e-3)))
    s = dataclasses.replace(s, ε_f=s.ε_f + 1e-6 * np.cos(np.arange(len(s.ε_f))))
    o = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    kw = dict(η_min=0.0, η_init=5e-2, maxiter=200, tol=1e-7, ω_acoustic=1e-6, method="auto")
    dev = epsx.Scgd0Device(s, ctx)
    ref = dev.loop(o, batched=True, **kw)
    assert ref["converged"].all() and ref["info"]["fm_method"] == 1
    r = dev.loop(o, batched=True, precision=prec, **kw)
    assert r["converged"].all() and r["info"]["fm_method"] == 1 and r["info"]["n_low_precision"] >= 1
    scale = np.max(np.abs(ref["Σ"]))
    assert np.max(np.abs(r["Σ"] - ref["Σ"])) <= 50 * kw["tol"] + 1e-9 * scale
    assert np.max(np.abs(r["μ"] - ref["μ"])) <= 1e-6
