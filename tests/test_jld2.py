"""JLD2 files of the results (CPU only): save_scgd0 / load_scgd0 (src/scgd0.jl:375-413) and
save_scgd0_ladder / load_scgd0_ladder (src/ladder.jl:380-416).

No HDF5 library is in this image, so the files are pinned by: the checksum function against the
published lookup3 test values; the byte layout of the superblock and of an object header read back
field by field as the HDF5 format specification lays them out; the key names the reference's test
checks (test/test_scgd0_spectral_grid.jl:69-73); bitwise round trips of every field
(test/test_scgd0_holstein.jl:53-57, test/test_ladder_pb.jl:24-35); and corruption detection.
"""
import struct

import numpy as np
import pytest

from epsx import FrequencyGrid, SCGD0Result, graded_frequency_grid
from epsx.jld2 import (HEADER_LEN, JLD2_HEADER, load_scgd0, load_scgd0_ladder, lookup3, read_tree, save_scgd0,
                       save_scgd0_ladder, write_tree)
from epsx.ladder import SCGD0LadderResult


def test_lookup3_known_values():
    """The values printed by the driver of Bob Jenkins' lookup3.c for hashlittle."""
    assert lookup3(b"", 0) == 0xDEADBEEF
    assert lookup3(b"", 0xDEADBEEF) == 0xBD5B7DDE
    assert lookup3(b"Four score and seven years ago", 0) == 0x17770551
    assert lookup3(b"Four score and seven years ago", 1) == 0xCD628161


def scgd0_result(rng, nT=2, niter=5):
    ωs = graded_frequency_grid([(-1.4, -0.4, 0.005), (-4.0, 4.0, 0.02)])
    wsa = FrequencyGrid(ωs[0], ωs.dω_min, np.arange(1601))
    nw, ni = len(ωs), 7
    Σ = np.asfortranarray(rng.standard_normal((nw, ni, nT)) + 1j * rng.standard_normal((nw, ni, nT)))
    h = dict(err=np.asfortranarray(rng.random((nT, niter))), μ=np.asfortranarray(rng.random((nT, niter))),
             min_spectral_weight=np.asfortranarray(rng.random((nT, niter))), wall=rng.random(niter),
             iter_frozen=np.array([3, 0], dtype=np.int64)[:nT])
    h["err"][1, 3:] = np.nan
    return SCGD0Result(ωs, wsa, Σ, rng.random(nT) - 0.5, np.array([0.01, 0.02])[:nT], np.zeros(nT),
                       np.array([True, False])[:nT], h)


def test_scgd0_round_trip_and_keys(tmp_path):
    rng = np.random.default_rng(3)
    res = scgd0_result(rng)
    f = save_scgd0(str(tmp_path / "scgd0.jld2"), res)
    tree = read_tree(f)
    # each grid is a group keyed by the field names of FrequencyGrid (test/test_scgd0_spectral_grid.jl:69-73)
    for grid in ("ωs", "ωs_spectral"):
        assert sorted(tree[grid].keys()) == ["dω_unit", "offsets", "ω_origin"]
    assert sorted(tree["history"].keys()) == sorted(["err", "μ", "min_spectral_weight", "wall", "iter_frozen"])
    assert sorted(k for k in tree if k not in ("ωs", "ωs_spectral", "history")) == sorted(
        ["Σ", "μ", "Tlist", "nlist", "converged"])
    back = load_scgd0(f)
    for name in ("ωs", "ωs_spectral"):
        g, h = getattr(back, name), getattr(res, name)
        assert g.ω_origin == h.ω_origin and g.dω_unit == h.dω_unit and g.dω_min == h.dω_min
        assert np.array_equal(g.offsets, h.offsets) and g.offsets.dtype == np.int64
    for name in ("Σ", "μ", "Tlist", "nlist", "converged"):
        a, b = getattr(back, name), getattr(res, name)
        assert a.dtype == np.asarray(b).dtype and a.shape == np.asarray(b).shape and np.array_equal(a, b), name
    for k, v in res.history.items():
        assert np.array_equal(back.history[k], v, equal_nan=True) and back.history[k].shape == v.shape, k
    assert back.history["iter_frozen"].dtype == np.int64


def test_file_layout(tmp_path):
    """The JLD2 header block, then the version-2 superblock at byte 512 with the base address at the
    superblock, its checksum, and the root object header at the address it names."""
    f = write_tree(str(tmp_path / "t.jld2"), {"x": np.arange(6, dtype=np.float64).reshape(2, 3, order="F"),
                                               "s": np.float64(2.5), "g": {"b": np.array([True, False])}})
    buf = open(f, "rb").read()
    assert buf.startswith(JLD2_HEADER) and buf[len(JLD2_HEADER):HEADER_LEN] == bytes(HEADER_LEN - len(JLD2_HEADER))
    sb = buf[HEADER_LEN:HEADER_LEN + 48]
    assert sb[:8] == b"\x89HDF\r\n\x1a\n" and sb[8:12] == bytes([2, 8, 8, 0])
    base, ext, eof, root = struct.unpack_from("<QQQQ", sb, 12)
    assert base == HEADER_LEN and ext == 0xFFFFFFFFFFFFFFFF and base + eof == len(buf)
    assert struct.unpack_from("<I", sb, 44)[0] == lookup3(sb[:44])
    assert buf[base + root:base + root + 6] == b"OHDR\x02\x02"
    tree = read_tree(f)
    # Julia size (2, 3) is stored with the HDF5 dimensions (3, 2), column-major data
    assert tree["x"].shape == (2, 3) and np.array_equal(tree["x"], np.arange(6.0).reshape(2, 3, order="F"))
    assert tree["s"] == 2.5 and np.ndim(tree["s"]) == 0
    assert tree["g"]["b"].dtype == np.bool_ and list(tree["g"]["b"]) == [True, False]
    # the data of x: 0..5 in column-major order, one contiguous run
    assert np.arange(6, dtype="<f8").tobytes() in buf


def test_corruption_is_detected(tmp_path):
    f = write_tree(str(tmp_path / "t.jld2"), {"x": np.ones(4)})
    buf = bytearray(open(f, "rb").read())
    root = struct.unpack_from("<Q", buf, HEADER_LEN + 36)[0]
    bad = bytearray(buf)
    bad[HEADER_LEN + root + 10] ^= 0xFF
    open(f, "wb").write(bytes(bad))
    with pytest.raises(ValueError, match="checksum"):
        read_tree(f)
    bad = bytearray(buf)
    bad[HEADER_LEN + 20] ^= 0x01
    open(f, "wb").write(bytes(bad))
    with pytest.raises(ValueError, match="checksum"):
        read_tree(f)


@pytest.mark.parametrize("store_vertex", [False, True])
def test_ladder_round_trip(tmp_path, store_vertex):
    rng = np.random.default_rng(5)
    nO, nT, nw, ni = 3, 2, 11, 4
    niter = np.array([[2, 0], [1, 3], [4, 2]], dtype=np.int64)
    hist = [[dict(norms=rng.random(niter[a, b] + 1), σ_iter=[rng.random((3, 3)) for _ in range(niter[a, b])])
             for b in range(nT)] for a in range(nO)]
    J = (np.asfortranarray(rng.standard_normal((nw, 4, 3, 2, ni, nO, nT)) + 1j) if store_vertex else None)
    res = SCGD0LadderResult(np.array([0.0, 0.1, 0.2]), np.array([0.01, 0.02]), rng.random(nT),
                            np.asfortranarray(rng.random((3, 3, nO, nT))), np.asfortranarray(rng.random((3, 3, nO, nT))),
                            np.asfortranarray(rng.random((nO, nT)) > 0.5), niter, hist, 0.001, J)
    f = save_scgd0_ladder(str(tmp_path / f"ladder_{store_vertex}.jld2"), res)
    tree = read_tree(f)
    assert sorted(tree["history"]["norms"].keys()) == sorted(f"{a}_{b}" for a in (1, 2, 3) for b in (1, 2))
    assert ("J" in tree) == store_vertex
    back = load_scgd0_ladder(f)
    for name in ("Ωs", "Tlist", "μ", "σ_bubble", "σ_ladder", "converged", "niter"):
        a, b = getattr(back, name), getattr(res, name)
        assert a.dtype == b.dtype and np.array_equal(a, b), name
    assert back.dω_unit == res.dω_unit
    assert (back.J is None) == (not store_vertex)
    if store_vertex:
        assert back.J.shape == J.shape and np.array_equal(back.J, J)
    for a in range(nO):
        for b in range(nT):
            assert np.array_equal(back.history[a][b]["norms"], hist[a][b]["norms"])
            assert len(back.history[a][b]["σ_iter"]) == niter[a, b]
            for m1, m2 in zip(back.history[a][b]["σ_iter"], hist[a][b]["σ_iter"]):
                assert np.array_equal(m1, m2)
