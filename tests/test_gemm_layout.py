"""Index bounds of the buffer-load FP64 GEMM (`dgemm_buf_kernel`, epspectral.jl_amd/csrc/epsx_gemm.hip),
checked on the CPU for every tile shape the autotuner launches (CPU only).

For every thread of a workgroup the LDS writes of the staged tiles, the MFMA operand reads and the
output rows / columns the accumulators map to are enumerated with the kernel's index formulas:
every LDS index lies inside the dynamic allocation the host sizes for the shape, every output
element of the tile is written by exactly one (thread, accumulator, element), and the staged k rows
and columns cover the tile exactly once.  Global reads go through buffer descriptors, whose range
check returns 0 for offsets past the operand, so they cannot fault; the output stores are guarded
by gi < m and gj < n.
"""
import itertools

import pytest

GKT = 16
SHAPES = [(64, 64, 2, 2), (128, 64, 2, 2), (64, 128, 2, 2), (128, 128, 2, 2), (256, 128, 4, 2), (256, 32, 4, 1),
          (128, 32, 4, 1)]


@pytest.mark.parametrize("bm,bn,wmw,wnw", SHAPES)
def test_buffer_gemm_indices_in_bounds(bm, bn, wmw, wnw):
    T = 64 * wmw * wnw
    PA, PB = bm + 16, bn + 2
    TM, TN = bm // wmw // 16, bn // wnw // 16
    QA, QB = bm * GKT // T, GKT * bn // T
    assert T % bm == 0 and T % GKT == 0 and QA * T == bm * GKT and QB * T == GKT * bn
    KA, CB = T // bm, T // GKT
    lds_doubles = 2 * GKT * PA + 2 * GKT * PB          # the host's dynamic LDS, in doubles
    assert 8 * lds_doubles <= 160 * 1024
    a_seen, b_seen = set(), set()
    for buf, tid, q in itertools.product(range(2), range(T), range(max(QA, QB))):
        rA, kA, kB, cB = tid % bm, tid // bm, tid % GKT, tid // GKT
        if q < QA:
            k = kA + KA * q
            idx = buf * GKT * PA + k * PA + rA                      # wA[buf GKT PA + KA q PA]
            assert 0 <= idx < 2 * GKT * PA and k < GKT
            a_seen.add((buf, k, rA))
        if q < QB:
            col = cB + CB * q
            idx = 2 * GKT * PA + buf * GKT * PB + kB * PB + col     # Bs after As
            assert 2 * GKT * PA <= idx < lds_doubles and col < bn
            b_seen.add((buf, kB, col))
    assert len(a_seen) == 2 * GKT * bm and len(b_seen) == 2 * GKT * bn   # each staged element once
    out = {}
    for tid in range(T):
        lane, wave = tid % 64, tid // 64
        wm, wn = wave % wmw, wave // wmw
        mb, nb = wm * (bm // wmw), wn * (bn // wnw)
        for cur, ks in itertools.product(range(2), range(0, GKT, 4)):
            ra = cur * GKT * PA + (lane >> 4) * PA + mb + (lane & 15) + ks * PA
            rb = 2 * GKT * PA + cur * GKT * PB + (lane >> 4) * PB + nb + (lane & 15) + ks * PB
            for a in range(TM):
                assert 0 <= ra + 16 * a < 2 * GKT * PA and mb + 16 * a + 15 < bm
            for b in range(TN):
                assert 2 * GKT * PA <= rb + 16 * b < lds_doubles and nb + 16 * b + 15 < bn
        for a, b, e in itertools.product(range(TM), range(TN), range(4)):
            i, j = mb + 16 * a + 4 * e + (lane >> 4), nb + 16 * b + (lane & 15)
            assert (i, j) not in out
            out[(i, j)] = tid
    assert len(out) == bm * bn and all(0 <= i < bm and 0 <= j < bn for i, j in out)


@pytest.mark.parametrize("kt", [8, 16])
def test_register_gemm_fragments(kt):
    """dgemm_reg_kernel (256 x 32 tiles, 4 waves of 64 rows, operands loaded straight into MFMA
    fragments): for every wave, k-step and lane, the A and B elements its buffer offsets address
    are those of the MFMA operand layout (A[i][k] of a 16x4 block in lane i + 16 k, B[k][j] in
    lane j + 16 k); over a k-tile each wave reads its 64 rows and the tile's 32 columns exactly
    once per k; the accumulators cover the tile exactly once."""
    lda, ldb = 1275, 8199
    TM, TN, KQ = 4, 2, kt // 4
    for wave in range(4):
        row0 = 64 * wave
        seen_a, seen_b = set(), set()
        for lane in range(64):
            voA = (lane & 15) + lda * (lane >> 4)
            voB = (lane >> 4) + ldb * (lane & 15)
            for q in range(KQ):
                kofs = 4 * q
                for a in range(TM):
                    off = voA + 16 * a + lda * kofs
                    i, k = off % lda, off // lda
                    assert (i, k) == (16 * a + (lane & 15), 4 * q + (lane >> 4))
                    seen_a.add((row0 + i, k))
                for b in range(TN):
                    off = voB + kofs + 16 * ldb * b
                    k, j = off % ldb, off // ldb
                    assert (k, j) == (4 * q + (lane >> 4), 16 * b + (lane & 15))
                    seen_b.add((k, j))
        assert seen_a == {(row0 + i, k) for i in range(64) for k in range(kt)}
        assert seen_b == {(k, j) for k in range(kt) for j in range(32)}
    out = {}
    for wave, lane in itertools.product(range(4), range(64)):
        for a, b, e in itertools.product(range(TM), range(TN), range(4)):
            i, j = 64 * wave + 16 * a + 4 * e + (lane >> 4), 16 * b + (lane & 15)
            assert (i, j) not in out
            out[(i, j)] = (wave, lane)
    assert len(out) == 256 * 32
    # the launched pipeline (8-deep k-tiles, 4 slots): at most 63 loads in flight per wave
    if kt == 8:
        assert (4 - 1) * KQ * (TM + TN) <= 63


@pytest.mark.parametrize("nkt", [0, 1, 2, 3, 7, 8, 257, 1025])
def test_register_gemm_two_waves_along_k(nkt):
    """dgemm_reg_kernel with WK = 2 (8 waves: 4 along m, 2 along k): the two waves of a row block
    take the k-tiles r = wk + 2 j, j < nj(wk), which cover the unit's tiles exactly once; the odd
    wave's accumulators go to the static LDS array xs[4 * 4 * 2 * 4 * 64] at indices inside it,
    each slot written once, and the even wave of the same rows reads back exactly the slots its
    partner wrote for the same (accumulator, element, lane)."""
    WK, TM, TN = 2, 4, 2
    tiles = []
    for wk in range(WK):
        nj = (nkt - wk + WK - 1) // WK if nkt > wk else 0
        tiles += [wk + WK * j for j in range(nj)]
    assert sorted(tiles) == list(range(nkt))
    size = 4 * TM * TN * 4 * 64
    written = {}
    for wm, lane, a, b, e in itertools.product(range(4), range(64), range(TM), range(TN), range(4)):
        idx = wm * (TM * TN * 4 * 64) + lane + ((a * TN + b) * 4 + e) * 64
        assert 0 <= idx < size and idx not in written
        written[idx] = (wm, lane, a, b, e)
    assert len(written) == size


@pytest.mark.parametrize("WM,WN,WK,TM,TN", [(4, 1, 2, 4, 2), (2, 1, 4, 8, 2), (2, 2, 1, 4, 4)])
def test_pair_gemm_fragments_and_reduction(WM, WN, WK, TM, TN):
    """dgemm_pair_kernel (tiles of 16 TM WM x 16 TN WN, WM x WN waves of TM x TN accumulators, WK
    along k, 16-byte loads): for every wave and lane the two doubles of each A load are the rows
    32 pr + 2 r + {0, 1} of the wave's block at k = kb + 2 g + step, which the MFMA takes as row r of
    accumulators 2 pr and 2 pr + 1 in lane slot g; the two doubles of each B load (one 16-byte or two
    8-byte loads at the same offsets) are k = kb + 2 g + {0, 1} of column 16 b + r of the wave's
    columns, one per step; over an 8-deep block each wave reads its rows and columns exactly once
    per k; the 16-byte loads are aligned for even pitches; the accumulators of the WM x WN waves
    cover the tile exactly once; the LDS halving rounds stay inside the allocation the host sizes
    and pair each writer with its reader."""
    lda, ldb = 1276, 8200
    NP, BM, BN = TM // 2, 16 * TM * WM, 16 * TN * WN
    for wm, wn in itertools.product(range(WM), range(WN)):
        row0, col0 = 16 * TM * wm, 16 * TN * wn
        seen_a, seen_b = set(), set()
        for lane in range(64):
            r, g = lane & 15, lane >> 4
            voA = 2 * r + lda * (2 * g)
            voB = 2 * g + ldb * r
            assert voA % 2 == 0 and voB % 2 == 0        # 16-byte aligned (8-byte elements)
            for st in range(2):
                for pr in range(NP):
                    off = voA + 32 * pr + lda * st
                    assert off % 2 == 0
                    for h in range(2):
                        i, k = (off + h) % lda, (off + h) // lda
                        assert (i, k) == (32 * pr + 2 * r + h, 2 * g + st)
                        seen_a.add((row0 + i, k))
            for b in range(TN):
                off = voB + 16 * ldb * b
                for h in range(2):
                    k, j = (off + h) % ldb, (off + h) // ldb
                    assert (k, j) == (2 * g + h, 16 * b + r)
                    seen_b.add((k, col0 + j))
        assert seen_a == {(row0 + i, k) for i in range(16 * TM) for k in range(8)}
        assert seen_b == {(k, col0 + j) for k in range(8) for j in range(16 * TN)}
    out = {}
    for wm, wn, lane in itertools.product(range(WM), range(WN), range(64)):
        for a, b, e in itertools.product(range(TM), range(TN), range(4)):
            i = 16 * TM * wm + 32 * (a >> 1) + 2 * (4 * e + (lane >> 4)) + (a & 1)
            j = 16 * TN * wn + 16 * b + (lane & 15)
            assert (i, j) not in out
            out[(i, j)] = (wm, wn, lane)
    assert len(out) == BM * BN and all(0 <= i < BM and 0 <= j < BN for i, j in out)
    # the launched pipelines: at most 63 loads in flight per wave (B as two 8-byte loads at most)
    KP = {(4, 1, 2): 4, (2, 1, 4): 2, (2, 2, 1): 3}[(WM, WN, WK)]
    assert (KP - 1) * (2 * NP + 2 * TN) <= 63
    # the LDS halving rounds
    if WK == 1:
        return
    PER = TM * TN * 4 * 64
    lds = (WK // 2) * WM * WN * PER
    assert 8 * lds <= 160 * 1024
    h = WK // 2
    while h >= 1:
        written = {}
        for wk, wm, wn, lane in itertools.product(range(h, 2 * h), range(WM), range(WN), range(64)):
            for a, b, e in itertools.product(range(TM), range(TN), range(4)):
                idx = ((wk - h) * WM * WN + wm + WM * wn) * PER + lane + ((a * TN + b) * 4 + e) * 64
                assert 0 <= idx < lds and idx not in written
                written[idx] = (wk, wm, wn, lane, a, b, e)
        for wk, wm, wn, lane in itertools.product(range(h), range(WM), range(WN), range(64)):
            for a, b, e in itertools.product(range(TM), range(TN), range(4)):
                idx = (wk * WM * WN + wm + WM * wn) * PER + lane + ((a * TN + b) * 4 + e) * 64
                assert written[idx] == (wk + h, wm, wn, lane, a, b, e)
        h //= 2


@pytest.mark.parametrize("gm", [2, 3])
def test_grouped_tile_order_covers_every_tile_once(gm):
    """unit_of_block with gm > 1 (the thin 16-byte-load tiles): the units of an instance, taken in
    groups of gm tile rows (row fastest, then column; a last smaller group for the remaining rows),
    cover every (tile row, tile column) exactly once, for tile grids smaller and larger than gm."""
    def tile(q, mt, nt):
        full = mt // gm * gm
        if q < full * nt:
            gi, w = divmod(q, gm * nt)
            return gi * gm + w % gm, w // gm
        rem = mt - full
        w = q - full * nt
        return full + w % rem, w // rem
    for mt, nt in ((5, 25), (10, 25), (4, 7), (1, 3), (3, 3), (2, 5), (7, 1)):
        seen = [tile(q, mt, nt) for q in range(mt * nt)]
        assert sorted(seen) == [(a, b) for a in range(mt) for b in range(nt)]
