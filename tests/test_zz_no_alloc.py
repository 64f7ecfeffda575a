"""epsx_kramers_kronig_noalloc (device buffers, the fixed workspace) against epsx_kramers_kronig
(host buffers, the library's allocations), through the C ABI, on two calls in a row, the second
on a smaller problem; and the size limit.  FP64: 1e-13 relative, the bar of every single-kernel
FP64 check (the same tile shape and k split give the same bits).

Device buffers come from hipMalloc of the HIP runtime libepsx itself links (so both see one
runtime), and are freed at the end.
"""
import ctypes as C

import numpy as np
import pytest

import epsx
from epsx import _lib
from epsx.grids import FrequencyGrid

pytestmark = pytest.mark.gpu

_lib.lib()
HIP = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")   # the runtime libepsx.so links (its rpath)
H2D, D2H = 1, 2


def hip(rc):
    assert rc == 0, f"HIP error {rc}"


class DeviceArray:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        self.p = C.c_void_p()
        hip(HIP.hipMalloc(C.byref(self.p), C.c_size_t(self.nbytes)))
        hip(HIP.hipMemcpy(self.p, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), H2D))

    def get(self):
        hip(HIP.hipDeviceSynchronize())
        out = np.empty_like(self.host)
        hip(HIP.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(self.nbytes), D2H))
        return out

    def free(self):
        hip(HIP.hipFree(self.p))


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


def noalloc(ctx, g, zT):
    """ys = K zs with zs column-major [n, ncol], passed as its transpose zT [ncol, n] in C order."""
    n, ncol = len(g), zT.shape[0]
    ws = DeviceArray(np.asarray(g, dtype=np.float64))
    z = DeviceArray(zT)
    y = DeviceArray(np.full_like(zT, np.nan))
    try:
        ctx.check(_lib.lib().epsx_kramers_kronig_noalloc(ctx.handle, n, ws.p, ncol, z.p, y.p, None))
        return y.get().T
    finally:
        for a in (ws, z, y):
            a.free()


def test_noalloc_matches_allocating_entry_twice(ctx):
    rng = np.random.default_rng(11)
    cases = [FrequencyGrid.from_nodes(np.concatenate([np.arange(-600, -300, 4), np.arange(-300, 301),
                                                      np.arange(304, 601, 4)]) * 1e-3),
             FrequencyGrid(-1.0, 0.01, np.arange(301))]
    held0 = ctx.device_memory()[0]
    for g, ncol in zip(cases, (257, 40)):
        n = len(g)
        zT = rng.standard_normal((ncol, n))
        ref = epsx.kramers_kronig(g, np.asfortranarray(zT.T), ctx)
        held = ctx.device_memory()[0]
        got = noalloc(ctx, g, zT)
        assert ctx.device_memory()[0] == held     # nothing allocated by the call
        assert np.all(np.isfinite(got))
        assert rel(got, ref) <= 1e-13
    assert held0 <= ctx.device_memory()[0]
    # beyond the workspace: a clear error, no work
    big = FrequencyGrid(0.0, 1e-3, np.arange(2049))
    with pytest.raises(epsx.ArgumentError, match="exceed the fixed workspace"):
        noalloc(ctx, big, np.ones((1, 2049)))
