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
