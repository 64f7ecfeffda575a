"""The ladder (vertex-corrected Kubo conductivity) set-up: the Python side of src/ladder.jl.

`setup_ladder` mirrors `_setup_ladder` (src/ladder.jl:112-158): the rotations R[:, :, f] of
the vertex (identity without symmetry), the inner states grouped by representative, the
hat-basis shift tables on the grid's own step, the quadrature weights and the velocities.
The device loop lives in libepsx (`LadderDevice`).
"""
import dataclasses
import math

import numpy as np

from ._lib import ArgumentError
from .grids import FrequencyGrid, trapezoid_weights, vertex_shift_tables
from .scgd0 import Scgd0Setup


@dataclasses.dataclass
class LadderSetup:
    scgd0: Scgd0Setup
    R: np.ndarray                # (3, 3, nf) rotation of each inner state's vertex
    v: np.ndarray                # (3, ni) band velocities of the outer states
    inner_states_by_star: np.ndarray   # 1-based, sortperm(f_to_i)
    inner_star_ptr: np.ndarray         # 1-based, length ni + 1
    nshift: int
    dω_unit: float
    ωs_weights: np.ndarray
    klo: np.ndarray
    wlo: np.ndarray
    whi: np.ndarray

    # the reference's ladder NamedTuple fields, read through to the scGD0 setup
    def __getattr__(self, k):
        if k in ("scgd0",):
            raise AttributeError(k)
        return getattr(self.scgd0, k)


def setup_ladder(setup: Scgd0Setup, R=None, v=None):
    """`_setup_ladder` (src/ladder.jl:112-158) on a scGD0 setup."""
    nm, ni, nf = setup.dims
    ωs = setup.ωs
    if not isinstance(ωs, FrequencyGrid):
        raise ArgumentError("the ladder needs ωs as a FrequencyGrid")
    if R is None:
        R = np.zeros((3, 3, nf), order="F")
        for a in range(3):
            R[a, a, :] = 1.0
    R = np.asfortranarray(R, dtype=np.float64)
    if R.shape != (3, 3, nf):
        raise ArgumentError(f"R must have size (3, 3, nf) = {(3, 3, nf)}")
    v = np.zeros((3, ni), order="F") if v is None else np.asfortranarray(v, dtype=np.float64)
    f_to_i = np.asarray(setup.f_to_i)
    by_star = np.argsort(f_to_i, kind="stable") + 1
    sizes = np.bincount(f_to_i - 1, minlength=ni)
    ptr = np.concatenate([[1], 1 + np.cumsum(sizes)]).astype(np.int64)
    h = ωs.dω_unit
    nshift = int(math.ceil(np.max(setup.ωph) / h)) + 2
    klo, wlo, whi = vertex_shift_tables(ωs, nshift)
    return LadderSetup(scgd0=setup, R=R, v=v, inner_states_by_star=by_star, inner_star_ptr=ptr, nshift=nshift,
                       dω_unit=float(h), ωs_weights=trapezoid_weights(ωs), klo=klo, wlo=wlo, whi=whi)
