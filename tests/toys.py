"""Small synthetic scGD0 setups for the parity tests (test data, not product code).

`dispersive_setup` is a toy with everything the Einstein-mode fixtures leave constant:
* phonon frequencies ωph[ν, q] that vary with q, one entry below ω_acoustic;
* a q map iq_kk[k_i, k_f] that is not symmetric (and, by default, not square);
* outer and inner k maps iks_i, iks_f that differ;
* per-state weights that differ, and a graded output grid.

With `identity = True` the inner states are the outer states (f -> i the identity), the form of
the reference's v2 test oracle (test/test_scgd0_pb.jl:12-67).  Otherwise the inner states are
twice as many, mapped onto the outer states by a fixed pseudo-random map, with energies that
differ from their representatives' (so the star method's precondition fails).
"""
import numpy as np

from epsx.grids import FrequencyGrid, uniform_spectral_grid
from epsx.models import bare_mu
from epsx.scgd0 import Scgd0Setup


def toy_grid():
    """A graded grid: 0.02 steps within ±0.4, 0.04 steps out to ±1.2 (81 nodes)."""
    nodes = np.concatenate([np.arange(-30, -10) * 0.04, np.arange(-20, 21) * 0.02, np.arange(11, 31) * 0.04])
    return FrequencyGrid.from_nodes(nodes)


def dispersive_setup(n=10, nm=2, nq=7, nki=5, nkf=6, identity=True, T0=0.03, ωs=None):
    """See the module docstring.  Energies in [-0.6, 0.6]; Σ w_i = 1; one electron per cell at
    μ_start = the bare μ at T0."""
    s = np.arange(n)
    ε_i = 0.6 * np.sin(1.7 * s + 0.3)
    w_i = 0.5 + 0.4 * np.cos(2.3 * s) ** 2
    w_i = w_i / w_i.sum()
    ν = np.arange(nm)[:, None]
    q = np.arange(nq)[None, :]
    ωph = 0.04 + 0.025 * ν + 0.013 * q + 0.004 * q * q * (ν + 1)
    ωph[0, 0] = 0.0                     # an acoustic entry: skipped before g2 is read
    a = np.arange(nki)[:, None]
    b = np.arange(nkf)[None, :]
    iq_kk = ((3 * a + 5 * b + a * b) % nq + 1).astype(np.int32)
    iks_i = (s % nki + 1).astype(np.int32)
    if identity:
        nf = n
        f_to_i = s + 1
        ε_f, w_f = ε_i.copy(), w_i.copy()
        iks_f = ((2 * s + 1) % nkf + 1).astype(np.int32)
    else:
        nf = 2 * n
        f = np.arange(nf)
        f_to_i = (7 * f + 3) % n + 1
        ε_f = ε_i[f_to_i - 1] + 0.01 * np.cos(1.1 * f)
        w_f = np.full(nf, 1.0 / nf)
        iks_f = ((3 * f + 2) % nkf + 1).astype(np.int32)
    r = np.arange(nm * n * nf).reshape((nm, n, nf), order="F")
    g2 = 0.002 * (1.0 + 0.8 * np.sin(0.37 * r) ** 2)
    ωs = toy_grid() if ωs is None else ωs
    μ0 = bare_mu(ε_i, w_i, T0, 1.0)
    return Scgd0Setup(
        g2=np.asfortranarray(g2), ωph=np.asfortranarray(ωph), iq_kk=np.asfortranarray(iq_kk),
        iks_i=iks_i, iks_f=iks_f, ε_i=ε_i, w_i=w_i, ε_f=ε_f, w_f=w_f,
        f_to_i=np.asarray(f_to_i, dtype=np.int32), W=float(w_i.sum()), nstates_base=0.0, multiplets=[],
        ωs=ωs, ωs_spectral=uniform_spectral_grid(ωs), μ_start=np.array([μ0]))


def pair_frequencies(s):
    """ωq[ν, i, f] = ωph[ν, iq_kk[iks_i[i], iks_f[f]]] by a plain loop (test/test_scgd0_pb.jl:80-81)."""
    nm, ni, nf = s.g2.shape
    out = np.empty((nm, ni, nf))
    for i in range(ni):
        for f in range(nf):
            qq = s.iq_kk[s.iks_i[i] - 1, s.iks_f[f] - 1] - 1
            for nu in range(nm):
                out[nu, i, f] = s.ωph[nu, qq]
    return out
