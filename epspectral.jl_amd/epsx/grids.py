"""Frequency grids and grid maps: the Python side of src/frequency_grids.jl.

`FrequencyGrid` mirrors the reference type (src/frequency_grids.jl:119-171): integer offsets on
a commensurate step, read-only, with the same constructor checks and error messages.
`graded_frequency_grid` mirrors :261-294.  The grid maps that feed the device
(`cell_search_table`, `linear_interpolation_table`, `trapezoid_weights`,
`vertex_shift_tables`) are computed by libepsx's host code, and `kramers_kronig_matrix` /
`kramers_kronig` on the device.
"""
import math

import numpy as np

from . import _lib
from ._lib import ArgumentError, ptr


def _jl(x):
    """A Float64 printed as Julia prints it (shortest round trip, `1.0e-12`, `0.5`)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    r = repr(x)
    if "e" in r:
        m, e = r.split("e")
        if "." not in m:
            m += ".0"
        return f"{m}e{int(e)}"
    if "." not in r:
        r += ".0"
    return r


class FrequencyGrid:
    """`ωs[j] = ω_origin + offsets[j] * dω_unit` with strictly increasing integer offsets
    (src/frequency_grids.jl:119-141).  `dω_min = dω_unit * minimum(diff(offsets))`.

    Behaves as a read-only vector of its nodes (`len`, indexing, `np.asarray`)."""

    __slots__ = ("ω_origin", "dω_unit", "offsets", "dω_min", "_nodes")

    def __init__(self, ω_origin, dω_unit, offsets):
        ω_origin, dω_unit = float(ω_origin), float(dω_unit)
        if not (math.isfinite(ω_origin) and math.isfinite(dω_unit) and dω_unit > 0):
            raise ArgumentError("ω_origin must be finite and dω_unit finite and positive, got "
                                f"{_jl(ω_origin)}, {_jl(dω_unit)}")
        off = np.array(offsets, dtype=np.int64).copy()
        if off.ndim != 1 or len(off) < 2:
            raise ArgumentError("a FrequencyGrid needs at least two nodes")
        gaps = np.diff(off)
        if not np.all(gaps > 0):
            k = int(np.argmax(gaps <= 0)) + 1
            raise ArgumentError(f"offsets must be strictly increasing, got offsets[{k}] >= offsets[{k + 1}]")
        off.setflags(write=False)
        object.__setattr__(self, "ω_origin", ω_origin)
        object.__setattr__(self, "dω_unit", dω_unit)
        object.__setattr__(self, "offsets", off)
        object.__setattr__(self, "dω_min", dω_unit * float(gaps.min()))
        nodes = ω_origin + off.astype(np.float64) * dω_unit
        nodes.setflags(write=False)
        object.__setattr__(self, "_nodes", nodes)

    def __setattr__(self, k, v):
        raise AttributeError("FrequencyGrid is read-only")

    @classmethod
    def from_nodes(cls, ωs, atol=1e-8):
        """`FrequencyGrid(ωs::AbstractVector; atol)` (src/frequency_grids.jl:143-171): the
        largest commensurate step `span / round(span / (dω_min / k))`, k <= 64."""
        ωs = np.asarray(ωs, dtype=np.float64)
        if isinstance(ωs, FrequencyGrid):
            ωs = np.asarray(ωs)
        n = len(ωs)
        if n < 2:
            raise ArgumentError("ωs must have at least two points")
        gaps = np.diff(ωs)
        nan = np.isnan(gaps)
        jmin = int(np.argmax(nan)) if nan.any() else int(np.argmin(gaps))
        dω_min = gaps[jmin]
        span = ωs[-1] - ωs[0]
        if not (dω_min > 1e-8 * span):
            raise ArgumentError(
                "ωs must be finite and strictly increasing with spacings above 1e-8 of its span; "
                f"ωs[{jmin + 1}] = {_jl(ωs[jmin])} and ωs[{jmin + 2}] = {_jl(ωs[jmin + 1])} are "
                f"{_jl(dω_min)} apart")
        best = (math.inf, 0, dω_min)
        for k in range(1, 65):
            unit = span / np.rint(span / (dω_min / k))
            p = (ωs - ωs[0]) / unit
            dev = np.abs(p - np.rint(p))
            j = int(np.argmax(np.isnan(dev))) if np.isnan(dev).any() else int(np.argmax(dev))
            deviation = dev[j]
            if deviation <= atol:
                return cls(ωs[0], unit, np.rint(p).astype(np.int64))
            if deviation < best[0]:
                best = (deviation, j, unit)
        dev, j, unit = best
        raise ArgumentError(
            "ωs is not commensurate: no step minimum(diff(ωs)) / k with k <= 64 puts every node "
            f"within {_jl(atol)} steps of ωs[1] + dω_unit ℤ; at the closest, dω_unit = {_jl(unit)}, "
            f"the node ωs[{j + 1}] = {_jl(ωs[j])} is {_jl(dev)} steps off. Build the grid with "
            "`graded_frequency_grid`, or move the node onto the grid.")

    # vector behaviour
    def __len__(self):
        return len(self.offsets)

    def __getitem__(self, k):
        return self._nodes[k]

    def __iter__(self):
        return iter(self._nodes)

    def __array__(self, dtype=None, copy=None):
        a = self._nodes
        return a.astype(dtype) if dtype is not None else a.copy()

    @property
    def nodes(self):
        return self._nodes

    def __eq__(self, other):
        return np.array_equal(np.asarray(self), np.asarray(other))

    def __repr__(self):
        return (f"FrequencyGrid(ω_origin={_jl(self.ω_origin)}, dω_unit={_jl(self.dω_unit)}, "
                f"{len(self)} nodes)")


def graded_frequency_grid(windows):
    """A grid whose step grows outward from nested windows `[(wmin, wmax, dw), ...]`
    (src/frequency_grids.jl:261-294): origin = centre of window 1, offsets on a dw_1 lattice,
    outer windows continued outward in steps dw_l / dw_1."""
    windows = [tuple(float(x) for x in w) for w in windows]
    if not windows:
        raise ArgumentError("windows must not be empty")
    wmins = np.array([w[0] for w in windows])
    wmaxs = np.array([w[1] for w in windows])
    dws = np.array([w[2] for w in windows])
    if not (np.all(dws > 0) and np.all(wmins < wmaxs)):
        raise ArgumentError(f"every window needs dw > 0 and wmin < wmax, got {windows}")
    if not (np.all(np.diff(wmins) <= 0) and np.all(np.diff(wmaxs) >= 0)):
        raise ArgumentError("windows must be nested from the smallest to the largest (wmin "
                            f"non-increasing, wmax non-decreasing), got {windows}")
    ω_origin, dw1 = (wmins[0] + wmaxs[0]) / 2, dws[0]
    tol = 1e-8
    u = (wmins - ω_origin) / dw1
    v = (wmaxs - ω_origin) / dw1
    ps = []
    for l in range(len(windows)):
        ratio = dws[l] / dw1
        if not abs(ratio - np.rint(ratio)) <= tol:
            raise ArgumentError(f"dw[{l + 1}] = {_jl(dws[l])} is not an integer multiple of dw[1] = "
                                f"{_jl(dw1)} (ratio {_jl(ratio)})")
        r = int(np.rint(ratio))
        if l == 0:
            ps.extend(range(math.ceil(u[0] - tol), math.floor(v[0] + tol) + 1))
            if len(ps) < 2:
                raise ArgumentError("the finest window must hold at least two nodes, got "
                                    f"{len(ps)} for {windows[0]}")
        else:
            p_lo, p_hi = min(ps), max(ps)
            ps.extend(range(p_hi + r, math.floor(v[l] + tol) + 1, r))
            ps.extend(range(p_lo - r, math.ceil(u[l] - tol) - 1, -r))
    return FrequencyGrid(ω_origin, dw1, sorted(ps))


def uniform_spectral_grid(ωs):
    """The default `ωs_spectral` of run_scgd0 (src/scgd0.jl:89-92): the uniform grid on the span
    of ωs with step ωs.dω_min."""
    # length(range(a, b; step = s)) = floor((b - a) / s) + 1, with Julia's tolerance for a
    # quotient a few ulps below an integer
    span = ωs[-1] - ωs[0]
    n = int(math.floor(span / ωs.dω_min * (1 + 1e-15))) + 1
    return FrequencyGrid(ωs[0], ωs.dω_min, np.arange(n))


def _nodes(ωs):
    return np.ascontiguousarray(np.asarray(ωs, dtype=np.float64))


def trapezoid_weights(ωs):
    """`trapezoid_weights` (src/frequency_grids.jl:53-54), by libepsx's host code."""
    x = _nodes(ωs)
    W = np.empty_like(x)
    rc = _lib.lib().epsx_trapezoid_weights(None, len(x), ptr(x), ptr(W))
    if rc:
        raise ArgumentError("trapezoid_weights: bad arguments")
    return W


def linear_interpolation_table(ωs, xs):
    """`linear_interpolation_table` (src/frequency_grids.jl:65-84): (j::int32 1-based, t)."""
    x = _nodes(ωs)
    q = np.ascontiguousarray(np.asarray(xs, dtype=np.float64))
    j = np.empty(len(q), dtype=np.int32)
    t = np.empty(len(q))
    rc = _lib.lib().epsx_linear_interpolation_table(None, len(x), ptr(x), len(q), ptr(q), ptr(j), ptr(t))
    if rc:
        raise ArgumentError("linear_interpolation_table: bad arguments")
    return j, t


def apply_interpolation_table(ys, j, t):
    """`apply_interpolation_table!` (src/frequency_grids.jl:93-96) on the host:
    Y[r, i] = (1 - t[r]) ys[j[r], i] + t[r] ys[j[r] + 1, i]."""
    ys = np.asarray(ys)
    j0 = np.asarray(j, dtype=np.int64) - 1
    t = np.asarray(t)
    tt = t.reshape((-1,) + (1,) * (ys.ndim - 1))
    return (1 - tt) * ys[j0] + tt * ys[j0 + 1]


def cell_search_table(ωs):
    """`cell_search_table` (src/frequency_grids.jl:184-196): (jtab::int32, x0, inv_h)."""
    if not isinstance(ωs, FrequencyGrid):
        raise ArgumentError("cell_search_table takes a FrequencyGrid")
    x = _nodes(ωs)
    L = _lib.lib()
    n = np.zeros(1, dtype=np.int64)
    x0 = np.zeros(1)
    ih = np.zeros(1)
    rc = L.epsx_cell_search_table(None, len(x), ptr(x), ωs.dω_min, None, 0, ptr(n), ptr(x0), ptr(ih))
    if rc:
        nb = (x[-1] - x[0]) / ((1 - 1e-6) * ωs.dω_min)
        if math.trunc(min(nb, 2.0**40)) + 1 > 2**24:
            raise ArgumentError(
                f"cell search table of {math.trunc(min(nb, 2.0**40)) + 1} entries for a grid with minimum "
                f"spacing {_jl(ωs.dω_min)} over a range {_jl(x[-1] - x[0])}; use a grid without a "
                "near-coincident pair of nodes")
        raise ArgumentError("two grid nodes fall in one bucket")
    jtab = np.empty(int(n[0]), dtype=np.int32)
    rc = L.epsx_cell_search_table(None, len(x), ptr(x), ωs.dω_min, ptr(jtab), len(jtab), ptr(n),
                                  ptr(x0), ptr(ih))
    if rc:
        raise ArgumentError("cell_search_table failed")
    return jtab, float(x0[0]), float(ih[0])


def vertex_shift_tables(ωs, nshift):
    """`vertex_shift_tables` (src/ladder_kernel.jl:158-176): klo (int32, 1-based), wlo, whi of
    size (nω, nshift, 2) in column-major order."""
    if not isinstance(ωs, FrequencyGrid):
        raise ArgumentError("vertex_shift_tables takes a FrequencyGrid")
    nw = len(ωs)
    off = np.ascontiguousarray(ωs.offsets)
    klo = np.empty((nw, nshift, 2), dtype=np.int32, order="F")
    wlo = np.empty((nw, nshift, 2), order="F")
    whi = np.empty((nw, nshift, 2), order="F")
    rc = _lib.lib().epsx_vertex_shift_tables(None, nw, ptr(off), int(nshift), ptr(klo), ptr(wlo), ptr(whi))
    if rc:
        raise ArgumentError("vertex_shift_tables: bad arguments")
    return klo, wlo, whi


def kramers_kronig_matrix(ωs, ctx=None):
    """`kramers_kronig_matrix` (src/frequency_grids.jl:30-44), built on the device."""
    ctx = ctx or _lib.default_context()
    x = _nodes(ωs)
    K = np.empty((len(x), len(x)), order="F")
    ctx.check(_lib.lib().epsx_kramers_kronig_matrix(ctx.handle, len(x), ptr(x), ptr(K)))
    return K


def kramers_kronig(ωs, zs, ctx=None):
    """`kramers_kronig(ωs, zs)` (src/frequency_grids.jl:227): Re part of the retarded function
    whose Im part is zs, K * zs on the device."""
    ctx = ctx or _lib.default_context()
    x = _nodes(ωs)
    z = np.asfortranarray(np.asarray(zs, dtype=np.float64))
    vec = z.ndim == 1
    z2 = z.reshape(len(x), -1, order="F")
    y = np.empty_like(z2, order="F")
    ctx.check(_lib.lib().epsx_kramers_kronig(ctx.handle, len(x), ptr(x), z2.shape[1], ptr(z2), ptr(y)))
    return y.reshape(-1) if vec else y.reshape(z.shape, order="F")
