"""run_scgd0's argument checks and defaults (src/scgd0.jl:81-113) on the host (CPU only).

Every check raises before anything reaches the device, so these run without a GPU; the texts are
the reference's `ArgumentError` / `@warn` messages.
"""
import dataclasses
import re
import warnings

import numpy as np
import pytest

import epsx
from epsx import ArgumentError, FrequencyGrid, Occupation
from epsx.grids import uniform_spectral_grid
from epsx.models import holstein_1d_setup
from epsx.scgd0 import check_run_arguments


def setup():
    ωs = FrequencyGrid.from_nodes(np.arange(-200, 201) * 0.02)
    return holstein_1d_setup(16, 1.0, 0.2, 0.2, ωs, [0.0])


def occ(**kw):
    kw.setdefault("Tlist", [0.1])
    kw.setdefault("nlist", 0.0)
    kw.setdefault("nelec", 1.0)
    return Occupation(**kw)


def raises(msg, **kw):
    s = setup()
    args = dict(occ=occ())
    args.update(kw)
    o = args.pop("occ")
    with pytest.raises(ArgumentError, match=re.escape(msg)):
        epsx.run_scgd0(s, o, **args)


def test_occupation_type_checks():
    """src/scgd0.jl:83-86."""
    raises("run_scgd0 uses the Fermi-Dirac occupation only, got occ.occ_type = GaussianSmearing",
           occ=occ(occ_type="GaussianSmearing"))
    raises("occ.type must be :Metal or :Semiconductor, got Insulator", occ=occ(type="Insulator"))


def test_grid_type_check():
    """src/scgd0.jl:87-88: ωs must be a FrequencyGrid."""
    s = setup()
    with pytest.raises(ArgumentError, match=r"ωs must be a FrequencyGrid\{Float64\}, got a ndarray; build one "
                                            r"with `graded_frequency_grid` or `FrequencyGrid\(ωs::AbstractVector\)`"):
        check_run_arguments(s, occ(), np.asarray(s.ωs), None, 0.0, None, 20, False)


def test_spectral_grid_checks_and_default():
    """src/scgd0.jl:89-106: the default uniform grid at dω_min over the span of ωs; a given grid
    must be a FrequencyGrid covering ωs to within dω_min; one past ωs is accepted with a warning."""
    s = setup()
    ωs = s.ωs
    g, η = check_run_arguments(s, occ(), ωs, None, 0.0, None, 20, False)
    assert isinstance(g, FrequencyGrid) and len(g) == 401 and g[0] == ωs[0] and g.dω_unit == ωs.dω_min
    assert np.array_equal(η, [0.0])
    raises("ωs_spectral must be a FrequencyGrid{Float64} or nothing, got a ndarray",
           ωs_spectral=np.asarray(ωs))
    short = FrequencyGrid(ωs[0] + 0.1, 0.01, np.arange(100))
    raises(f"ωs_spectral spans [{short[0]}, {short[-1]}], which does not cover ωs = [{ωs[0]}, {ωs[-1]}] "
           "to within ωs.dω_min", ωs_spectral=short)
    # within dω_min of the ends: accepted silently
    near = FrequencyGrid(ωs[0] + 0.01, 0.01, np.arange(799))
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        check_run_arguments(s, occ(), ωs, near, 0.0, None, 20, False)
    wide = FrequencyGrid(ωs[0] - 1.0, 0.02, np.arange(501))
    with pytest.warns(UserWarning, match=re.escape(f"ωs_spectral spans [{wide[0]}, {wide[-1]}], past ωs = "
                                                   f"[{ωs[0]}, {ωs[-1]}]; Σ is held at its end values outside ωs")):
        check_run_arguments(s, occ(), ωs, wide, 0.0, None, 20, False)


def test_fix_mu_eta_min_gpu_tile_maxiter_checks():
    """src/scgd0.jl:107-113."""
    raises("fix_μ = true keeps μ at occ.μlist, which is not given", fix_μ=True)
    raises("η_min must be a scalar or have one entry per temperature (1), got 2", η_min=[0.0, 0.1])
    raises("gpu_tile must be nothing or a positive number of inner states, got 0", gpu_tile=0)
    raises("maxiter must be at least 1, got 0", maxiter=0)
    # a per-temperature η_min of the right length and a scalar are both accepted
    s = setup()
    o2 = occ(Tlist=[0.1, 0.2])
    _, η = check_run_arguments(s, o2, s.ωs, None, [0.0, 0.03], 5, 1, False)
    assert np.array_equal(η, [0.0, 0.03])
    _, η = check_run_arguments(s, o2, s.ωs, None, 0.02, None, 1, False)
    assert np.array_equal(η, [0.02, 0.02])


def test_sigma_init_size_check():
    """src/scgd0.jl:209-210, checked on the host before the device is touched."""
    s = setup()
    dev = epsx.Scgd0Device.__new__(epsx.Scgd0Device)   # no upload: the check comes first
    dev.setup, dev.ctx = s, None
    bad = np.zeros((len(s.ωs), 3, 1), dtype=complex)
    with pytest.raises(ArgumentError, match=re.escape(f"Σ_init must have size (nω, n_i, nT) = "
                                                      f"({len(s.ωs)}, 9, 1), got (401, 3, 1)")):
        dev.loop(occ(), η_min=0.0, η_init=0.05, maxiter=2, tol=1e-6, Σ_init=bad)
