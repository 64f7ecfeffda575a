"""GPU parity of the warm-start network (libepsx through its C ABI) against the float64 NumPy
oracle (oracle/net_oracle.py), and the warm start's gain on scGD0.

Tolerances (relative, in the 2-norm over all outputs / all gradient entries): FP32 MFMA path
1e-5 (forward) and 1e-4 (gradient); BF16 operands with FP32 accumulation 3e-2 and 6e-2.
"""
import dataclasses
import warnings

import numpy as np
import pytest

import epsx
import net_oracle
from epsx.learn import (WarmStartNet, g0d0_and_converged, make_training_set, predict_sigma, random_t2g_model,
                        run_warm_started)
from epsx.scgd0 import OMEGA_ACOUSTIC

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _data(cfg, n, L, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, cfg[2], L)).astype(np.float32)
    t = rng.normal(scale=0.5, size=(n, cfg[4], L)).astype(np.float32)
    return x, t


@pytest.mark.parametrize("layers,kernel,hidden,L,n", [(3, 5, 16, 70, 5), (6, 9, 64, 1024, 3)])
def test_forward_and_gradient_match_oracle(ctx, layers, kernel, hidden, L, n):
    net = WarmStartNet(ctx, L, n_layers=layers, kernel=kernel, c_hidden=hidden, seed=11)
    p = net.params
    cfg = net.cfg
    rng = np.random.default_rng(2)
    # a trained-like net: larger last-layer weights than the initialisation
    p = p + rng.normal(scale=0.02, size=p.shape).astype(np.float32)
    net.params = p
    x, t = _data(cfg, n, L, 4)
    y_ref = net_oracle.forward(p.astype(np.float64), x, cfg)
    y32 = net.forward(x, "fp32")
    ybf = net.forward(x, "bf16")
    assert rel(y32, y_ref) <= 1e-5
    assert rel(ybf, y_ref) <= 3e-2
    loss_ref, g_ref = net_oracle.loss_and_grad(p.astype(np.float64), x, t, cfg)
    l32, g32 = net.gradient(x, t, "fp32")
    lbf, gbf = net.gradient(x, t, "bf16")
    assert abs(l32 - loss_ref) <= 1e-5 * loss_ref
    assert rel(g32, g_ref) <= 1e-4
    assert abs(lbf - loss_ref) <= 3e-2 * loss_ref
    assert rel(gbf, g_ref) <= 6e-2
    # bitwise reproducible
    assert np.array_equal(net.gradient(x, t, "bf16")[1], gbf)


def test_training_reduces_loss_and_is_reproducible(ctx):
    L = 128
    net = WarmStartNet(ctx, L, n_layers=4, kernel=7, c_hidden=32, seed=3)
    p0 = net.params
    rng = np.random.default_rng(8)
    x = rng.normal(size=(96, 5, L)).astype(np.float32)
    # a learnable target: a smoothed combination of the first two input channels
    k = np.exp(-0.5 * (np.arange(-3, 4) / 1.5) ** 2)
    k /= k.sum()
    sm = np.apply_along_axis(lambda r: np.convolve(r, k, mode="same"), 2, x[:, :2, :])
    t = (0.5 * sm[:, ::-1, :] - 0.2 * x[:, 2:4, :]).astype(np.float32)
    l0, _ = net.gradient(x, t, "fp32")
    # trained straight after creation: the optimiser state starts at zero
    losses = net.train(x, t, epochs=30, batch=16, lr=2e-3, precision="bf16", seed=1)
    assert np.all(np.isfinite(losses))
    l1, _ = net.gradient(x, t, "fp32")
    assert losses[-1] < 0.2 * losses[0] and l1 < 0.2 * l0
    p1 = net.params
    net.params = p0
    net.train(x, t, epochs=30, batch=16, lr=2e-3, precision="bf16", seed=1)
    assert np.array_equal(net.params, p1)
    # a fresh network of the same seed trains to the same weights as a reset one
    fresh = WarmStartNet(ctx, L, n_layers=4, kernel=7, c_hidden=32, seed=3)
    fresh.train(x, t, epochs=30, batch=16, lr=2e-3, precision="bf16", seed=1)
    assert np.array_equal(fresh.params, p1)


def test_warm_start_cuts_iterations(ctx):
    """Trained on 12 random models (nk = 8, 512 uniform nodes), the warm start converges held-out
    models in fewer iterations than the plain start, to the same fixed point; a corrupted network
    falls back to the plain start."""
    L = 512
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        x, t = make_training_set(ctx, 12, nk=8, length=L, seed=100)
        net = WarmStartNet(ctx, L, seed=0)
        losses = net.train(x, t, epochs=60, batch=32, lr=1e-3, precision="bf16", seed=0)
        assert losses[-1] < 0.5 * losses[0]
        rng = np.random.default_rng(999)
        gains = []
        for _ in range(3):
            setup, meta = random_t2g_model(rng, nk=8, length=L)
            dev = epsx.Scgd0Device(setup, ctx)
            occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
            _, rc = g0d0_and_converged(dev, occ)
            n_plain = rc["history"]["err"].shape[1]
            w = run_warm_started(dev, setup, occ, net)
            assert w["converged"].all() and rc["converged"].all()
            check_judge(dev, setup, occ, net, w)
            # the same fixed point to within the tolerance scale
            assert np.max(np.abs(w["Σ"] - rc["Σ"])) < 1e-4
            gains.append((n_plain, w["iterations"], bool(w["fallback"][0])))
        print("plain vs warm-started iterations (fallback):", gains)
        assert any(wi < npl for npl, wi, fb in gains), gains
        # a corrupted network: huge output bias -> the prediction is rejected
        p = net.params
        p[-2:] = 50.0
        net.params = p
        setup, meta = random_t2g_model(rng, nk=8, length=L)
        dev = epsx.Scgd0Device(setup, ctx)
        occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
        w = run_warm_started(dev, setup, occ, net)
        assert w["fallback"].all() and w["converged"].all()
        check_judge(dev, setup, occ, net, w)


def check_judge(dev, setup, occ, net, w):
    """The device warm start against the same steps run from the host: the first iteration from
    -iη_init, the prediction of `predict_sigma` on the host features, the second iteration of the
    predicted and the plain start as one batch of 2 nT; the device run keeps the slot with the
    smaller residual, and its second history entry is that residual."""
    kw = dict(η_min=0.0, η_init=5e-3, ω_acoustic=OMEGA_ACOUSTIC, method="auto", batched=True)
    nT = len(occ)
    r1 = dev.loop(occ, maxiter=1, tol=0.0, **kw)
    Σp = np.stack([predict_sigma(net, r1["Σ"][:, :, t], setup.ωs, setup.ε_i, r1["μ"][t], occ.Tlist[t], "bf16")
                   for t in range(nT)], axis=2)
    occ2 = dataclasses.replace(occ, Tlist=np.concatenate([occ.Tlist, occ.Tlist]),
                               nlist=np.concatenate([occ.nlist, occ.nlist]), μlist=None)
    try:
        r2 = dev.loop(occ2, maxiter=1, tol=0.0, Σ_init=np.concatenate([Σp, r1["Σ"]], axis=2),
                      μ_start=np.concatenate([r1["μ"], r1["μ"]]), **kw)
        ep, eq = r2["history"]["err"][:nT, 0], r2["history"]["err"][nT:, 0]
    except epsx.ArgumentError:   # no μ root for the prediction
        r2 = dev.loop(occ, maxiter=1, tol=0.0, Σ_init=r1["Σ"], μ_start=r1["μ"], **kw)
        ep, eq = np.full(nT, np.inf), r2["history"]["err"][:, 0]
    assert np.array_equal(~w["fallback"], ep < eq), (ep, eq)
    np.testing.assert_allclose(w["history"]["err"][:, 0], r1["history"]["err"][:, 0], rtol=1e-12)
    np.testing.assert_allclose(w["history"]["err"][:, 1], np.minimum(ep, eq), rtol=1e-5)


def test_c5_config_network_held_out_24(ctx):
    """The configuration network (6 layers, 64 channels, width 9, 1024 nodes) trained on random
    12³ models; a held-out 24³ model run from the device warm start reaches the plain run's fixed
    point, and a corrupted network falls back to the plain start with the same result."""
    L = 1024
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        x, t = make_training_set(ctx, 16, nk=12, length=L, seed=31, states_per_model=96)
        net = WarmStartNet(ctx, L, seed=1)
        assert net.nparam > 1e5
        losses = net.train(x, t, epochs=30, batch=64, lr=1e-3, precision="bf16", seed=0)
        assert losses[-1] < 0.5 * losses[0]
        rng = np.random.default_rng(4242)
        done = 0
        for _ in range(4):
            setup, meta = random_t2g_model(rng, nk=24, length=L)
            dev = epsx.Scgd0Device(setup, ctx)
            occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
            try:
                _, rc = g0d0_and_converged(dev, occ)
            except epsx.ArgumentError:
                continue   # no μ root for this random model
            if not rc["converged"].all():
                continue   # the plain iteration itself does not converge: no fixed point to compare
            w = run_warm_started(dev, setup, occ, net)
            assert w["converged"].all()
            assert np.max(np.abs(w["Σ"] - rc["Σ"])) < 1e-5
            check_judge(dev, setup, occ, net, w)
            print("24³ held-out: plain", rc["history"]["err"].shape[1], "warm", w["iterations"],
                  "kept", (~w["fallback"]).tolist())
            # with the BF16 / FP32 precision levels: the warm slot leaves the BF16 level at its
            # floor as the plain one does, and the run ends at the FP64 fixed point
            kw = dict(η_min=0.0, η_init=5e-3, ω_acoustic=OMEGA_ACOUSTIC, method="auto", batched=True)
            wl = dev.loop(occ, maxiter=300, tol=1e-6, precision="bf16", warm_start="bf16", **kw)
            pl = dev.loop(occ, maxiter=300, tol=1e-6, precision="bf16", **kw)
            assert wl["converged"].all() and pl["converged"].all()
            assert np.max(np.abs(wl["Σ"] - rc["Σ"])) < 1e-5
            n_wl, n_pl = wl["history"]["err"].shape[1], pl["history"]["err"].shape[1]
            print("24³ held-out, BF16 levels: plain", n_pl, "warm", n_wl, "kept", wl["info"]["warm_kept"].tolist())
            if wl["info"]["warm_kept"].all():
                assert n_wl <= n_pl + 5
            p = net.params
            bad = p.copy()
            bad[-2:] = 50.0
            net.params = bad
            wb = run_warm_started(dev, setup, occ, net)
            net.params = p
            assert wb["fallback"].all() and wb["converged"].all()
            assert np.max(np.abs(wb["Σ"] - rc["Σ"])) < 1e-5
            done += 1
            break
        assert done == 1, "no held-out 24³ model with a converging plain run"


def test_warm_start_argument_checks(ctx):
    """The warm start's preconditions: a network over the nω nodes, batched runs, no Σ_init,
    maxiter >= 3; each refused with an ArgumentError before anything runs."""
    rng = np.random.default_rng(5)
    setup, meta = random_t2g_model(rng, nk=8, length=128)
    dev = epsx.Scgd0Device(setup, ctx)
    occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    kw = dict(η_min=0.0, η_init=5e-3, ω_acoustic=OMEGA_ACOUSTIC, method="auto", tol=1e-6)
    WarmStartNet(ctx, 64, seed=0)   # a network over another length
    with pytest.raises(epsx.ArgumentError, match="nodes of ωs"):
        dev.loop(occ, maxiter=10, batched=True, warm_start="bf16", **kw)
    WarmStartNet(ctx, 128, seed=0)
    with pytest.raises(epsx.ArgumentError, match="batched"):
        dev.loop(occ, maxiter=10, batched=False, warm_start="bf16", **kw)
    with pytest.raises(epsx.ArgumentError, match="maxiter >= 3"):
        dev.loop(occ, maxiter=2, batched=True, warm_start="bf16", **kw)
    Σ0 = np.full((len(setup.ωs), setup.dims[1], len(occ)), -5e-3j)
    with pytest.raises(epsx.ArgumentError, match="Σ_init"):
        dev.loop(occ, maxiter=10, batched=True, warm_start="bf16", Σ_init=Σ0, **kw)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r = dev.loop(occ, maxiter=10, batched=True, warm_start="fp32", **kw)
    assert r["info"]["warm_kept"].shape == (len(occ),)
