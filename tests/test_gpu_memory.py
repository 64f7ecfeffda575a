"""The context's device memory: released temporaries kept for reuse, and the fallback that empties
them when an allocation would not fit otherwise (GPU, through the C ABI).

The allocation cap EPSX_ALLOC_CAP_MB (read when a context opens) makes an allocation fail as if the
device were full once the context would hold more than the cap, so the fallback is exercised
without filling the device.
"""
import dataclasses
import os
import warnings

import numpy as np
import pytest

import epsx
from epsx.grids import FrequencyGrid
from epsx.models import holstein_1d_setup

pytestmark = pytest.mark.gpu

OCC8 = epsx.Occupation(Tlist=np.linspace(0.1, 0.3, 8), nlist=0.0, nelec=1.0)
OCC1 = epsx.Occupation(Tlist=[0.1], nlist=0.0, nelec=1.0)
KW = dict(η_min=0.0, η_init=0.05, maxiter=2, tol=0.0, fix_μ=True, method="direct")


def holstein_modes(nk, nm):
    """The Holstein chain with its coupling split equally over nm identical modes (the same Σ as
    one mode, up to the order of the sum)."""
    ωs = FrequencyGrid.from_nodes(np.arange(-200, 201) * 0.02)
    s = holstein_1d_setup(nk, 1.0, 0.2, 0.2, ωs, [0.0] * 8)
    ni = s.g2.shape[1]
    return dataclasses.replace(s, g2=np.full((nm, ni, nk), 0.2 / nm, order="F"),
                               ωph=np.full((nm, nk), 0.2, order="F"))


def run(ctx, s, o):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return epsx.Scgd0Device(s, ctx).loop(o, **dict(KW, μ_start=[0.0] * len(o)))


def test_pool_is_emptied_when_an_allocation_needs_it():
    # A: large temporaries (513 outer states x 8 temperatures), small coupling;
    # B: large coupling (2048 modes), small temporaries (33 outer states, 1 temperature)
    A, B = holstein_modes(1024, 1), holstein_modes(64, 2048)
    c = epsx.Context(0)
    try:
        run(c, A, OCC8)
        h1, p1 = c.device_memory()
        epsx.Scgd0Device(B, c)
        h2, p2 = c.device_memory()
    finally:
        c.close()
    c = epsx.Context(0)
    try:
        rB = run(c, B, OCC1)
        hB, _ = c.device_memory()       # B's problem and all its run's temporaries
    finally:
        c.close()
    # the A run released its temporaries to the pool, and B's upload kept them (and released its
    # own set-up temporaries next to them)
    assert p1 > 0 and p2 >= p1
    # a cap that holds A's run and B's run on their own, but not B's upload next to A's pool
    lo = max(h1, hB)
    assert lo < h2
    cap_mb = ((lo + h2) // 2) >> 20
    assert lo >> 20 < cap_mb < h2 >> 20
    os.environ["EPSX_ALLOC_CAP_MB"] = str(cap_mb)
    try:
        c = epsx.Context(0)
    finally:
        del os.environ["EPSX_ALLOC_CAP_MB"]
    try:
        run(c, A, OCC8)
        assert c.device_memory() == (h1, p1)
        dev = epsx.Scgd0Device(B, c)
        held, pooled = c.device_memory()
        # uploaded by emptying the pool: what it keeps now is at most B's own set-up temporaries
        assert held <= cap_mb << 20 and pooled <= p2 - p1 < p1
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            r = dev.loop(OCC1, **dict(KW, μ_start=[0.0]))
        assert np.array_equal(r["Σ"], rB["Σ"])
        # the split coupling gives the one-mode Σ up to the order of the sum
        r1 = run(c, holstein_modes(64, 1), OCC1)
        assert np.linalg.norm(r["Σ"] - r1["Σ"]) <= 1e-12 * np.linalg.norm(r1["Σ"])
        # a coupling that does not fit even with the pool empty fails with the out-of-memory error
        with pytest.raises(epsx.DeviceError, match="device allocation"):
            epsx.Scgd0Device(holstein_modes(64, 2 * (cap_mb << 20) // (33 * 64 * 8) + 1), c)
    finally:
        c.close()
