"""The learned Σ_init warm start (no reference counterpart: the reference's hook is the Σ_init
argument of `_loop_scgd0`, src/scgd0.jl:66-67, :204-212).

A 1D convolutional network over the frequency grid predicts, for every outer state, the
converged self-energy from the G0D0 one (the first iteration from -iη_init), conditioned on
T, ε_i and μ.  The network runs in libepsx (hand-written HIP forward, backward and Adam
kernels, BF16 MFMA with FP32 master weights); this module builds its inputs and targets,
generates training sets from scGD0 runs of seeded random models, and runs scGD0 from the
prediction with a fallback: the plain iteration always runs to the same tol, and the warm start
is kept only if its first iteration leaves a smaller residual than the plain start's.

Features per outer state i (channels over the L grid points):
  0, 1: asinh(Re Σ_G0D0 / SIGMA_SCALE), asinh(Im Σ_G0D0 / SIGMA_SCALE) (the G0D0 self-energy of
        a narrow -iη_init start has spikes of tens of eV; asinh keeps them in range)
  2:    (ω - μ) / ENERGY_SCALE
  3:    (ε_i - μ) / ENERGY_SCALE (constant)
  4:    T / T_SCALE (constant)
Targets: Σ_converged / SIGMA_SCALE, real and imaginary parts (the converged self-energy is
smooth and of order 1 eV, unlike the correction from the spiky G0D0 one).
"""
import ctypes
import sys
import time
import warnings

import numpy as np

from . import _lib
from ._lib import ArgumentError

SIGMA_SCALE = 0.05    # eV
ENERGY_SCALE = 1.0    # eV
T_SCALE = 0.04        # eV (about 460 K)
N_FEATURES = 5


class WarmStartNet:
    """The network in a libepsx context: create, forward, gradient, train, parameters."""

    def __init__(self, ctx, length, n_layers=6, kernel=9, c_hidden=64, seed=0):
        self.ctx = ctx
        self.cfg = (n_layers, kernel, N_FEATURES, c_hidden, 2)
        self.length = int(length)
        c = _lib.NetConfig(ctypes.sizeof(_lib.NetConfig), n_layers, kernel, N_FEATURES, c_hidden, 2, 0, self.length)
        ctx.check(_lib.lib().epsx_net_create(ctx.handle, ctypes.addressof(c), ctypes.c_uint64(seed)))
        n = np.zeros(1, dtype=np.int64)
        ctx.check(_lib.lib().epsx_net_param_count(ctx.handle, _lib.ptr(n)))
        self.nparam = int(n[0])

    @property
    def params(self):
        p = np.empty(self.nparam, dtype=np.float32)
        self.ctx.check(_lib.lib().epsx_net_get_params(self.ctx.handle, _lib.ptr(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (self.nparam,):
            raise ArgumentError(f"expected {self.nparam} parameters, got {p.shape}")
        self.ctx.check(_lib.lib().epsx_net_set_params(self.ctx.handle, _lib.ptr(p)))

    def _check_x(self, x, nch):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 3 or x.shape[1] != nch or x.shape[2] != self.length:
            raise ArgumentError(f"samples must be [n, {nch}, {self.length}], got {x.shape}")
        return x

    def forward(self, x, precision="bf16"):
        x = self._check_x(x, N_FEATURES)
        y = np.empty((x.shape[0], 2, self.length), dtype=np.float32)
        self.ctx.check(_lib.lib().epsx_net_forward(self.ctx.handle, _lib.ptr(x), x.shape[0],
                                                   _lib.PRECISIONS[precision], _lib.ptr(y)))
        return y

    def gradient(self, x, t, precision="fp32"):
        x = self._check_x(x, N_FEATURES)
        t = self._check_x(t, 2)
        loss = np.zeros(1)
        g = np.empty(self.nparam, dtype=np.float32)
        self.ctx.check(_lib.lib().epsx_net_gradient(self.ctx.handle, _lib.ptr(x), _lib.ptr(t), x.shape[0],
                                                    _lib.PRECISIONS[precision], _lib.ptr(loss), _lib.ptr(g)))
        return float(loss[0]), g

    def train(self, x, t, epochs=20, batch=64, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0,
              precision="bf16", seed=0, verbosity=0):
        """Adam over shuffled mini-batches; returns the mean batch loss of every epoch."""
        x = self._check_x(x, N_FEATURES)
        t = self._check_x(t, 2)
        if len(x) != len(t):
            raise ArgumentError("inputs and targets differ in count")
        p = _lib.NetTrainParams(ctypes.sizeof(_lib.NetTrainParams), int(epochs), int(batch), float(lr), float(beta1),
                                float(beta2), float(eps), float(weight_decay), _lib.PRECISIONS[precision],
                                int(verbosity), int(seed))
        out = np.empty(max(int(epochs), 1))
        self.ctx.check(_lib.lib().epsx_net_train(self.ctx.handle, _lib.ptr(x), _lib.ptr(t), len(x),
                                                 ctypes.addressof(p), _lib.ptr(out)))
        return out[:epochs]

    def save(self, path):
        np.savez(path, params=self.params, cfg=np.array(self.cfg), length=self.length)

    def load(self, path):
        d = np.load(path)
        if tuple(d["cfg"]) != self.cfg or int(d["length"]) != self.length:
            raise ArgumentError("saved network has another architecture")
        self.params = d["params"]


def _by_states(fn, n):
    """fn(a, b) over slices [a, b) of n outer states, on up to 8 threads (NumPy releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    nt = max(1, min(8, n // 64))
    cuts = [n * k // nt for k in range(nt + 1)]
    if nt == 1:
        fn(0, n)
        return
    with ThreadPoolExecutor(nt) as ex:
        list(ex.map(lambda k: fn(cuts[k], cuts[k + 1]), range(nt)))


def features(Σ0, ωs, ε_i, μ, T):
    """Inputs [ni, 5, nω] of the network for the G0D0 self-energy Σ0 [nω, ni] at one T (asinh of
    Σ0 / SIGMA_SCALE in FP64, stored FP32, by slices of outer states on host threads)."""
    Σ0 = np.asarray(Σ0)
    ws = np.asarray(ωs, dtype=np.float64)
    nw, ni = Σ0.shape
    x = np.empty((ni, N_FEATURES, nw), dtype=np.float32)

    def part(a, b):
        St = np.ascontiguousarray(Σ0[:, a:b].T)
        x[a:b, 0, :] = np.arcsinh(St.real / SIGMA_SCALE)
        x[a:b, 1, :] = np.arcsinh(St.imag / SIGMA_SCALE)

    _by_states(part, ni)
    x[:, 2, :] = ((ws - μ) / ENERGY_SCALE)[None, :]
    x[:, 3, :] = ((np.asarray(ε_i) - μ) / ENERGY_SCALE)[:, None]
    x[:, 4, :] = T / T_SCALE
    return x


def targets(Σc):
    """Targets [ni, 2, nω]: the converged Σc [nω, ni], scaled."""
    d = np.asarray(Σc) / SIGMA_SCALE
    return np.ascontiguousarray(np.stack([d.real.T, d.imag.T], axis=1), dtype=np.float32)


def predict_sigma(net, Σ0, ωs, ε_i, μ, T, precision="bf16"):
    """Σ [nω, ni] predicted from the G0D0 Σ0 [nω, ni] at one temperature."""
    y = net.forward(features(Σ0, ωs, ε_i, μ, T), precision)
    ni, _, nw = y.shape
    out = np.empty((ni, nw), dtype=np.complex128)
    s = np.float32(SIGMA_SCALE)

    def part(a, b):
        # the scale applied in FP32, as to the network's FP32 outputs
        out[a:b].real = y[a:b, 0, :] * s
        out[a:b].imag = y[a:b, 1, :] * s

    _by_states(part, ni)
    return out.T


def random_t2g_model(rng, nk=12, length=1024, half_width=3.0, Tlist_K=None):
    """A seeded random three-band model: t ~ U(0.2, 0.35) eV (t' = 0.03 t / 0.26), median |g|²
    log-uniform in [5e-4, 5e-3] eV², T ~ U(50, 500) K (or the given temperatures), on `length`
    uniform ω nodes over bare μ ± half_width, or (length None) on the benchmark's graded grid."""
    from .models import t2g_setup
    t = rng.uniform(0.2, 0.35)
    g2m = float(np.exp(rng.uniform(np.log(5e-4), np.log(5e-3))))
    TK = rng.uniform(50.0, 500.0)
    seed = int(rng.integers(1 << 31))
    return t2g_setup(nk=nk, t=t, tp=0.03 * t / 0.26, g2_median=g2m, Tlist_K=(TK,) if Tlist_K is None else Tlist_K,
                     seed=seed, uniform_grid=None if length is None else (length, half_width))


def g0d0_and_converged(dev, occ, η_init=5e-3, tol=1e-6, maxiter=300, method="auto"):
    """The G0D0 self-energy (one iteration from -iη_init), its μ, and the converged plain run."""
    from .scgd0 import OMEGA_ACOUSTIC
    kw = dict(η_min=0.0, η_init=η_init, ω_acoustic=OMEGA_ACOUSTIC, method=method)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r1 = dev.loop(occ, maxiter=1, tol=0.0, **kw)
        rc = dev.loop(occ, maxiter=maxiter, tol=tol, **kw)
    return r1, rc


def make_training_set(ctx, n_models, nk=12, length=1024, seed=0, states_per_model=None, verbosity=0):
    """(x, t) training pairs from scGD0 runs of `n_models` random models (G0D0 -> converged)."""
    from .scgd0 import Occupation, Scgd0Device
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    for m in range(n_models):
        setup, meta = random_t2g_model(rng, nk=nk, length=length)
        dev = Scgd0Device(setup, ctx)
        occ = Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
        try:
            r1, rc = g0d0_and_converged(dev, occ)
        except ArgumentError:
            continue   # a model whose window holds no μ for the electron count: not a sample
        if not rc["converged"].all():
            continue
        Σ0, Σc = r1["Σ"][:, :, 0], rc["Σ"][:, :, 0]
        x = features(Σ0, setup.ωs, setup.ε_i, r1["μ"][0], meta["Tlist"][0])
        t = targets(Σc)
        if not (np.isfinite(x).all() and np.isfinite(t).all()):
            continue
        if states_per_model is not None and states_per_model < len(x):
            pick = rng.choice(len(x), size=states_per_model, replace=False)
            x, t = x[pick], t[pick]
        xs.append(x)
        ts.append(t)
        if verbosity:
            print(f"[learn] model {m + 1}/{n_models}: ni {len(setup.ε_i)}, plain iterations "
                  f"{rc['history']['err'].shape[1]}", file=sys.stderr, flush=True)
    return np.concatenate(xs), np.concatenate(ts)


def run_warm_started(dev, setup, occ, net, *, η_init=5e-3, tol=1e-6, maxiter=300, method="auto",
                     precision="bf16"):
    """scGD0 to tol from the network's prediction, with the fallback to the plain start, in one
    batched run on the device (`Scgd0Device.loop(..., warm_start=precision)`):

    1. the first iteration from -iη_init gives Σ_G0D0 and μ;
    2. the network (in the same context as `dev`) predicts Σ_p from Σ_G0D0 into a second set of
       slots; the second iteration runs both the plain iterate and Σ_p, and per temperature the
       slot with the smaller residual goes on (the plain start whenever the prediction does not
       cut the residual, or when no μ satisfies the electron count for it);
    3. the loop continues from the kept slot to tol.
    Returns dict(Σ, μ, converged, iterations, fallback[T], wall)."""
    from .scgd0 import OMEGA_ACOUSTIC
    if net.ctx is not dev.ctx:
        raise ArgumentError("the network and the scGD0 setup must live in the same context")
    kw = dict(η_min=0.0, η_init=η_init, ω_acoustic=OMEGA_ACOUSTIC, method=method, batched=True)
    t0 = time.perf_counter()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        rf = dev.loop(occ, maxiter=maxiter, tol=tol, warm_start=precision, **kw)
    wall = time.perf_counter() - t0
    n = rf["history"]["err"].shape[1]
    return dict(Σ=rf["Σ"], μ=rf["μ"], converged=rf["converged"], iterations=n,
                fallback=~rf["info"]["warm_kept"], history=rf["history"], wall=wall)