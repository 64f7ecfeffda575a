"""Frequency grids and grid maps (CPU): the oracle and the product's host code against the
reference's known answers (tests/golden/reference_known_answers.json, transcribed from
test/test_frequency_grids.jl), and the product's host tables bitwise against the oracle."""
import json
import math
import os

import numpy as np
import pytest

import epsx
import oracle
from epsx import ArgumentError, FrequencyGrid, graded_frequency_grid
from ref_numpy import pseudorandom, pseudorandom_complex, interp_flat

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))

SRVO3 = np.concatenate([np.arange(-350, -109, 20), np.arange(-99, 101), np.arange(120, 341, 20)]) / 1000.0  # Julia's correctly rounded ranges


def fl(x):
    return float(x) if not isinstance(x, str) else float(x.replace("Inf", "inf").replace("NaN", "nan"))


def both_from_nodes(ws, atol=1e-8):
    """(product grid, oracle (origin, unit, offsets))"""
    return FrequencyGrid.from_nodes(ws, atol), oracle.grid_from_nodes(ws, atol)


def test_frequency_grid_structure():
    g0 = GOLD["frequency_grid_basic"]
    g = FrequencyGrid(g0["origin"], g0["unit"], g0["offsets"])
    assert list(np.asarray(g)) == g0["nodes"] and g.dω_min == g0["dw_min"]
    assert len(g) == 4 and g[-1] == 1.75 and list(g[1:3]) == [0.5, 0.75]
    offs = [0, 1, 3]
    g = FrequencyGrid(0.0, 0.5, offs)
    offs[1] = 7
    assert list(g.offsets) == [0, 1, 3]
    with pytest.raises(AttributeError):
        g.dω_unit = 1.0
    for o, u, off in GOLD["frequency_grid_errors"]["cases"]:
        with pytest.raises(ArgumentError):
            FrequencyGrid(fl(o), fl(u), off)


def test_from_nodes_steps():
    g0 = GOLD["from_nodes_uniform"]
    ws = g0["start"] + g0["step"] * np.arange(g0["n"])
    g, (o, u, off) = both_from_nodes(ws)
    assert math.isclose(g.dω_unit, g0["unit"], rel_tol=1e-14) and g.ω_origin == 0.0
    assert list(g.offsets) == list(range(g0["n"]))
    assert (o, u, list(off)) == (g.ω_origin, g.dω_unit, list(g.offsets))
    g, (o, u, off) = both_from_nodes(SRVO3)
    assert math.isclose(g.dω_unit, 0.001, rel_tol=1e-12) and math.isclose(g.dω_min, 0.001, rel_tol=1e-12)
    # `≈` on vectors is norm-based
    assert np.linalg.norm(np.asarray(g) - SRVO3) <= 1e-14 * np.linalg.norm(SRVO3)
    assert (o, u) == (g.ω_origin, g.dω_unit) and np.array_equal(off, g.offsets)
    ws = np.concatenate([np.arange(0, 11) * 0.05, 0.52 + 0.02 * np.arange(15)])
    g, (o, u, off) = both_from_nodes(ws)
    assert math.isclose(g.dω_unit, 0.01, rel_tol=1e-12) and math.isclose(g.dω_min, 0.02, rel_tol=1e-12)
    assert u == g.dω_unit
    ws = np.concatenate([-0.5 + 0.05 * np.arange(9), -0.09 + 0.01 * np.arange(19), 0.1 + 0.05 * np.arange(9)])
    assert math.isclose(FrequencyGrid.from_nodes(ws).dω_unit, 0.01, rel_tol=1e-12)
    # a fine step far from zero, in Ry (test/test_frequency_grids.jl:41-49)
    eV = 1 / 13.605693122994
    c = 11.83 * eV
    ws = c + np.arange(-20000, 20001) * 1e-4 * eV
    for g in (FrequencyGrid.from_nodes(ws),
              graded_frequency_grid([(c - 0.1 * eV, c + 0.1 * eV, 1e-4 * eV), (c - eV, c + eV, 2e-3 * eV)])):
        assert math.isclose(g.dω_unit, 1e-4 * eV, rel_tol=1e-12)
        o, u, off = oracle.grid_from_nodes(np.asarray(g))
        assert math.isclose(u, 1e-4 * eV, rel_tol=1e-12)


def test_from_nodes_tolerance_and_errors():
    nodes = np.arange(11) * 0.1
    n9 = nodes.copy()
    n9[5] += 1e-9 * 0.1
    g = FrequencyGrid.from_nodes(n9)
    assert list(g.offsets) == list(range(11)) and abs(g[5] - n9[5]) <= 2e-8 * g.dω_unit
    assert list(oracle.grid_from_nodes(n9)[2]) == list(range(11))
    n7 = nodes.copy()
    n7[5] += 1e-7 * 0.1
    with pytest.raises(ArgumentError) as e:
        FrequencyGrid.from_nodes(n7)
    assert f"ωs[6] = {float(n7[5])!r}" in str(e.value) and "not commensurate" in str(e.value)
    with pytest.raises(oracle.OracleError) as eo:
        oracle.grid_from_nodes(n7)
    assert eo.value.args[:2] == (2, 6)
    assert list(FrequencyGrid.from_nodes(n7, atol=1e-6).offsets) == list(range(11))
    bad = [0.0, 1.0, 1.0 + math.sqrt(2) / 100]
    with pytest.raises(ArgumentError) as e:
        FrequencyGrid.from_nodes(bad)
    msg = str(e.value)
    assert "not commensurate" in msg and "graded_frequency_grid" in msg and "steps off" in msg
    for b in ([0.0, 0.5, 0.5, 1.0], [0.0, 1.0, 0.5], [0.0, 0.5, np.nextafter(0.5, 1), 1.0], [0.0, 1e-12, 0.5],
              [0.0, float("nan"), 1.0]):
        with pytest.raises(ArgumentError, match="apart"):
            FrequencyGrid.from_nodes(b)
        with pytest.raises(oracle.OracleError) as eo:
            oracle.grid_from_nodes(np.asarray(b))
        assert eo.value.args[0] == 1
    with pytest.raises(ArgumentError):
        FrequencyGrid.from_nodes([0.0])


def test_graded_grids():
    g0 = GOLD["graded_srvo3"]
    g = graded_frequency_grid(g0["windows"])
    assert g.ω_origin == g0["origin"] and g.dω_unit == g0["unit"] and g.dω_min == 0.001
    assert g.offsets[0] == g0["first_offset"] and g.offsets[-1] == g0["last_offset"] and len(g) == g0["n"]
    assert sum(1 for x in g if 0.4 <= x <= 0.6 + 1e-12) == g0["n_inner"]
    o, u, off = oracle.graded_grid(g0["windows"])
    assert (o, u) == (g.ω_origin, g.dω_unit) and np.array_equal(off, g.offsets)
    for key in ("graded_list_1", "graded_list_2"):
        gl = GOLD[key]
        g = graded_frequency_grid(gl["windows"])
        assert list(np.asarray(g)) == gl["nodes"]
        o, u, off = oracle.graded_grid(gl["windows"])
        assert list(o + off * u) == gl["nodes"]
        # every cell at most the step of the innermost window holding both ends
        for j, cell in enumerate(np.diff(np.asarray(g))):
            dw = min(w[2] for w in gl["windows"] if w[0] - 1e-12 <= g[j] and g[j + 1] <= w[1] + 1e-12)
            assert cell <= dw + 1e-12
    g3 = graded_frequency_grid(GOLD["graded_grid3"]["windows"])
    assert g3.dω_unit == 0.01 and g3.ω_origin == 0.05 and g3[-1] <= 1.0 + 1e-12 and 1.0 - g3[-1] < 0.03
    # the vector conversion recovers the step and offset differences
    for g in (graded_frequency_grid(GOLD["graded_srvo3"]["windows"]), g3,
              graded_frequency_grid(GOLD["graded_list_1"]["windows"])):
        gv = FrequencyGrid.from_nodes(np.asarray(g))
        assert math.isclose(gv.dω_unit, g.dω_unit, rel_tol=1e-12)
        assert np.array_equal(np.diff(gv.offsets), np.diff(g.offsets))
    assert np.allclose(np.asarray(graded_frequency_grid([(-1.0, 1.0, 0.25)])), np.arange(-4, 5) * 0.25)
    assert np.allclose(np.asarray(graded_frequency_grid([(0.0, 1.0, 0.2)])), GOLD["graded_single"]["odd_nodes"])
    for w in GOLD["graded_errors"]["cases"]:
        with pytest.raises(ArgumentError):
            graded_frequency_grid(w)
        with pytest.raises(oracle.OracleError):
            oracle.graded_grid(w if w else np.zeros((0, 3)))


def lookup_bisection(Sin, col, x, ws):
    n = len(ws)
    if not x > ws[0]:
        return Sin[0, col]
    if x >= ws[n - 1]:
        return Sin[n - 1, col]
    j = int(np.searchsorted(ws, x, side="right")) - 1
    t = (x - ws[j]) / (ws[j + 1] - ws[j])
    return (1 - t) * Sin[j, col] + t * Sin[j + 1, col]


@pytest.mark.parametrize("which", ["srvo3", "fine500"])
def test_cell_search_bitwise_bisection(which):
    if which == "srvo3":
        ws = np.asarray(FrequencyGrid.from_nodes(SRVO3))
        xs = np.linspace(ws[0] - 0.05, ws[-1] + 0.05, 997)
    else:
        ws = np.concatenate([[-1.0, -0.5], np.arange(-0.499, -0.4 + 1e-12, 1e-3), np.arange(0.1, 1.0 + 1e-9, 0.5)])
        ws = np.asarray(FrequencyGrid.from_nodes(ws))
        xs = np.linspace(-1.2, 1.2, 2001)
    g = FrequencyGrid.from_nodes(ws)
    Sin = pseudorandom_complex(len(ws), 2, seed=4)
    table = oracle.cell_search_table(ws, g.dω_min)
    # the product's host table is the oracle's, bitwise
    jt, x0, ih = epsx.cell_search_table(g)
    assert np.array_equal(jt, table[0]) and (x0, ih) == table[1:]
    pts = np.concatenate([xs, np.nextafter(ws, np.inf), np.nextafter(ws, -np.inf), [np.nan]])
    for col in range(2):
        for x in pts:
            a = lookup_bisection(Sin, col, x, ws)
            b = oracle.interp_linear_grid(Sin, col, x, ws, table)
            assert a == b or (a != a and b != b)
    if which == "fine500":
        assert len(table[0]) <= (ws[-1] - ws[0]) / 0.999e-3 + 2


def test_cell_search_refused():
    c = GOLD["cell_table_refused"]
    g = FrequencyGrid(c["origin"], c["unit"], c["offsets"])
    with pytest.raises(ArgumentError, match="cell search table"):
        epsx.cell_search_table(g)
    with pytest.raises(oracle.OracleError):
        oracle.cell_search_table(np.asarray(g), g.dω_min)


def test_interpolation_tables():
    g = FrequencyGrid.from_nodes(SRVO3)
    ws = np.asarray(g)
    wsa = np.asarray(epsx.uniform_spectral_grid(g))
    xs = np.concatenate([np.linspace(ws[0] - 0.05, ws[-1] + 0.05, 997), ws, wsa])
    j, t = oracle.linear_interpolation_table(ws, xs)
    jp, tp = epsx.linear_interpolation_table(g, xs)
    assert np.array_equal(j, jp) and np.array_equal(t, tp)
    assert j.min() >= 1 and j.max() <= len(ws) - 1 and t.min() >= 0 and t.max() <= 1
    y = pseudorandom_complex(len(ws), 3, seed=2)
    Y = oracle.apply_interpolation_table(y, j, t)
    Y_ref = np.stack([interp_flat(xs, ws, y[:, i]) for i in range(3)], axis=1)
    assert np.linalg.norm(Y - Y_ref) <= 1e-13 * np.linalg.norm(Y_ref)
    assert np.array_equal(Y[0], y[0]) and np.array_equal(Y[996], y[-1])
    assert np.array_equal(epsx.apply_interpolation_table(y, j, t), Y)
    # the table lookup of the cell search agrees with numpy.interp on the same points
    table = oracle.cell_search_table(ws, g.dω_min)
    S = pseudorandom_complex(len(ws), 3, seed=4)
    for col in range(3):
        err = max(abs(oracle.interp_linear_grid(S, col, x, ws, table) - interp_flat(np.array([x]), ws, S[:, col])[0])
                  for x in xs)
        assert err <= 1e-13 * np.abs(S).max()


def test_trapezoid_weights():
    ws = np.array([-1.0, -0.7, -0.6, 0.0, 0.5, 2.0])
    W = oracle.trapezoid_weights(ws)
    assert np.array_equal(W, epsx.trapezoid_weights(ws))
    assert abs(W.sum() - (ws[-1] - ws[0])) <= 1e-14
    assert math.isclose(W[0], (ws[1] - ws[0]) / 2) and math.isclose(W[-1], (ws[-1] - ws[-2]) / 2)
    assert abs(np.sum(W * ws) - (ws[-1] ** 2 - ws[0] ** 2) / 2) <= 1e-14


def test_kk_simpson_and_single_pole():
    from ref_numpy import simpson_pv_kk
    c = GOLD["kk_simpson_grid"]
    xg = np.array(c["xg"])
    zg = pseudorandom(len(xg), seed=c["seed"])
    y = oracle.kramers_kronig_matrix(xg) @ zg
    y_ref = simpson_pv_kk(xg, zg)
    assert np.linalg.norm(y - y_ref) <= c["tol"] * np.linalg.norm(y_ref)
    p = GOLD["kk_single_pole"]
    ws = np.arange(round((p["hi"] - p["lo"]) / p["step"]) + 1) * p["step"] + p["lo"]
    f = 1 / (ws - p["eps"] + 1j * p["eta"])
    K = oracle.kramers_kronig_matrix(ws)
    inds = np.abs(ws) < p["window"]
    assert np.max(np.abs((K @ f.imag)[inds] - f.real[inds])) < p["tol"]
    assert np.max(np.abs((-K @ f.real)[inds] - f.imag[inds])) < p["tol"]


def test_occupations():
    for e, T in ((0.3, 0.1), (-0.2, 0.05), (0.0, 0.3)):
        assert oracle.occ_fermion(e, T) == 1 / (math.exp(e / T) + 1)
        assert oracle.occ_boson(0.2, T) == 1 / math.expm1(0.2 / T)
    assert oracle.occ_boson(0.0, 0.1) == -0.5
    assert oracle.occ_fermion(0.1, 0.0) == 0.0 and oracle.occ_fermion(-0.1, 0.0) == 1.0 and oracle.occ_fermion(0.0, 0.0) == 0.5
    assert oracle.occ_boson(0.1, 0.0) == 0.0 and oracle.occ_boson(0.0, 0.0) == -0.5 and oracle.occ_boson(-0.1, 0.0) == -1.0
    assert math.isnan(oracle.occ_fermion(0.1, -1.0))
