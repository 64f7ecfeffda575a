"""epsx: the MI355X (gfx950) implementation of EPSpectral.jl's scGD0 hot path.

The Python layer mirrors the reference's API (FrequencyGrid, graded_frequency_grid,
kramers_kronig, run_scgd0 / SCGD0Result on a setup) and drives libepsx.so, the HIP library
behind the C ABI of include/epsx.h.  All compute runs in libepsx on the GPU; there is no CPU
fallback.
"""
from ._lib import ArgumentError, Context, DeviceError, default_context, lib
from .grids import (FrequencyGrid, apply_interpolation_table, cell_search_table,
                    graded_frequency_grid, kramers_kronig, kramers_kronig_matrix,
                    linear_interpolation_table, trapezoid_weights, uniform_spectral_grid,
                    vertex_shift_tables)
from .jld2 import load_scgd0, load_scgd0_ladder, save_scgd0, save_scgd0_ladder
from .ladder import (LadderDevice, LadderSetup, SCGD0LadderResult, run_scgd0_ladder, setup_ladder,
                     symmetrize)
from .scgd0 import (ELECTRON_DEGEN_CUTOFF, OMEGA_ACOUSTIC, Occupation, SCGD0Result, Scgd0Device,
                    Scgd0Setup, build_multiplets, loop_scgd0, run_scgd0)

__all__ = [
    "ArgumentError", "Context", "DeviceError", "default_context", "lib",
    "FrequencyGrid", "apply_interpolation_table", "cell_search_table", "graded_frequency_grid",
    "kramers_kronig", "kramers_kronig_matrix", "linear_interpolation_table", "trapezoid_weights",
    "uniform_spectral_grid", "vertex_shift_tables",
    "ELECTRON_DEGEN_CUTOFF", "OMEGA_ACOUSTIC", "Occupation", "SCGD0Result", "Scgd0Device",
    "Scgd0Setup", "build_multiplets", "loop_scgd0", "run_scgd0",
    "load_scgd0", "load_scgd0_ladder", "save_scgd0", "save_scgd0_ladder",
    "LadderDevice", "LadderSetup", "SCGD0LadderResult", "run_scgd0_ladder", "setup_ladder", "symmetrize",
]
