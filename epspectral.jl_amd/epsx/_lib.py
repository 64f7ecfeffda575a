"""ctypes bindings of libepsx.so (the C ABI declared in include/epsx.h).

The library is loaded from this directory.  There is no fallback: if libepsx.so is missing or
no GPU is present, using it raises, so nothing ever runs on a silent CPU path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libepsx.so")

EPSX_ABI_VERSION = 2
EPSX_OK, EPSX_ERR_ARG, EPSX_ERR_NO_MU_ROOT, EPSX_ERR_DEVICE, EPSX_ERR_OOM, EPSX_ERR_STATE = range(6)
FM_METHODS = {"direct": 0, "einstein": 1, "star": 2, "auto": 3}
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2}


class ArgumentError(ValueError):
    """Invalid input: the reference's `ArgumentError`."""


class DeviceError(RuntimeError):
    """A HIP runtime failure reported by libepsx."""


_lib = None

# (name, restype, argtypes) of every exported entry point of include/epsx.h
P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32
D = C.c_double
EXPORTS = [
    ("epsx_abi_version", C.c_int, []),
    ("epsx_open", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("epsx_close", None, [P]),
    ("epsx_last_error", C.c_char_p, [P]),
    ("epsx_last_timings", C.c_int, [P, P, P]),
    ("epsx_device_memory", C.c_int, [P, P, P]),
    ("epsx_cell_search_table", C.c_int, [P, I64, P, D, P, I64, P, P, P]),
    ("epsx_linear_interpolation_table", C.c_int, [P, I64, P, I64, P, P, P]),
    ("epsx_trapezoid_weights", C.c_int, [P, I64, P, P]),
    ("epsx_vertex_shift_tables", C.c_int, [P, I64, P, I64, P, P, P]),
    ("epsx_kramers_kronig_matrix", C.c_int, [P, I64, P, P]),
    ("epsx_kramers_kronig", C.c_int, [P, I64, P, I64, P, P]),
    ("epsx_kramers_kronig_noalloc", C.c_int, [P, I64, P, I64, P, P, P]),
    ("epsx_add_fan_migdal_imsigma", C.c_int,
     [P, I64, I64, I64, P, P, P, P, P, I64, D, D, P, I64, P, D, D, D, I32, P]),
    ("epsx_scgd0_set_problem", C.c_int, [P, P]),
    ("epsx_scgd0_run", C.c_int, [P, P, P]),
    ("epsx_ladder_set_problem", C.c_int, [P, P]),
    ("epsx_ladder_run", C.c_int, [P, P, P]),
    ("epsx_apply_ladder_kernel", C.c_int, [P, P, P, P, D, D, I32, P]),
    ("epsx_net_create", C.c_int, [P, P, C.c_uint64]),
    ("epsx_net_param_count", C.c_int, [P, P]),
    ("epsx_net_get_params", C.c_int, [P, P]),
    ("epsx_net_set_params", C.c_int, [P, P]),
    ("epsx_net_forward", C.c_int, [P, P, I64, I32, P]),
    ("epsx_net_gradient", C.c_int, [P, P, P, I64, I32, P, P]),
    ("epsx_net_train", C.c_int, [P, P, P, I64, P, P]),
]


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run epspectral.jl_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, res, args in EXPORTS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.epsx_abi_version() != EPSX_ABI_VERSION:
            raise ImportError("libepsx ABI version mismatch")
        _lib = L
    return _lib


class NetConfig(C.Structure):
    _fields_ = [("struct_size", I64), ("n_layers", I32), ("kernel", I32), ("c_in", I32), ("c_hidden", I32),
                ("c_out", I32), ("pad", I32), ("length", I64)]


class NetTrainParams(C.Structure):
    _fields_ = [("struct_size", I64), ("epochs", I64), ("batch", I64), ("lr", D), ("beta1", D), ("beta2", D),
                ("eps", D), ("weight_decay", D), ("precision", I32), ("verbosity", I32), ("seed", C.c_uint64)]


class Scgd0Problem(C.Structure):
    _fields_ = [
        ("struct_size", I64),
        ("nm", I64), ("ni", I64), ("nf", I64),
        ("nq", I64), ("nki", I64), ("nkf", I64),
        ("g2", P), ("wph", P), ("iq_kk", P), ("iks_i", P), ("iks_f", P),
        ("e_i", P), ("w_i", P), ("e_f", P), ("w_f", P), ("f_to_i", P),
        ("W", D), ("nstates_base", D),
        ("nw", I64), ("w_origin", D), ("dw_unit", D), ("w_offsets", P),
        ("nwa", I64), ("wa_origin", D), ("dwa_unit", D), ("wa_offsets", P),
        ("n_mult", I64), ("mult_ptr", P), ("mult_idx", P),
    ]


class Scgd0Params(C.Structure):
    _fields_ = [
        ("struct_size", I64),
        ("nT", I64),
        ("Tlist", P), ("nlist", P), ("mu_start", P), ("eta_min", P),
        ("nelec", D), ("spin_degeneracy", D),
        ("is_semiconductor", I32), ("fix_mu", I32),
        ("eta_init", D), ("tol", D), ("w_acoustic", D),
        ("maxiter", I64), ("f_tile", I64),
        ("fm_method", I32), ("precision", I32), ("batched", I32), ("verbosity", I32),
        ("Sigma_init", P),
        ("reuse_star", I32), ("warm_start", I32),
    ]


class Scgd0Outputs(C.Structure):
    _fields_ = [
        ("struct_size", I64),
        ("Sigma", P), ("mu", P), ("converged", P),
        ("hist_err", P), ("hist_mu", P), ("hist_minA", P), ("wall", P), ("iter_frozen", P),
        ("niter", I64),
        ("fm_method_used", I32), ("n_low_precision", I32),
        ("warm_kept", P),
    ]


class LadderProblem(C.Structure):
    _fields_ = [("struct_size", I64), ("R", P), ("v", P)]


class LadderParams(C.Structure):
    _fields_ = [
        ("struct_size", I64), ("nT", I64), ("nOmega", I64),
        ("Tlist", P), ("mu", P), ("Sigma", P), ("Omegas", P),
        ("solver", I32), ("precision", I32), ("rtol", D), ("maxiter", I64), ("w_acoustic", D),
        ("store_vertex", I32), ("verbosity", I32), ("batched", I32), ("pad", I32),
    ]


class LadderOutputs(C.Structure):
    _fields_ = [
        ("struct_size", I64), ("sigma_bubble", P), ("sigma_ladder", P), ("converged", P), ("niter", P),
        ("norms", P), ("sigma_iter", P), ("J", P),
    ]


def ptr(a):
    """Data pointer of a contiguous numpy array (None for None)."""
    if a is None:
        return None
    return a.ctypes.data


def f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def i32(a):
    return np.asfortranarray(a, dtype=np.int32)


def i64(a):
    return np.asfortranarray(a, dtype=np.int64)


def c128(a):
    return np.asfortranarray(a, dtype=np.complex128)


class Context:
    """A device context (one HIP stream) of libepsx: the `backend` of this implementation."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.epsx_open(device, C.byref(h))
        if rc != EPSX_OK:
            raise DeviceError(f"epsx_open(device={device}) failed with code {rc}: no usable GPU")
        self._h = h
        self.device = device

    @property
    def handle(self):
        if self._h is None:
            raise DeviceError("context is closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().epsx_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self):
        return lib().epsx_last_error(self.handle).decode("utf-8", "replace")

    def check(self, rc):
        if rc == EPSX_OK:
            return
        msg = self.last_error()
        if rc in (EPSX_ERR_ARG, EPSX_ERR_NO_MU_ROOT):
            raise ArgumentError(msg)
        raise DeviceError(f"libepsx error {rc}: {msg}")

    def device_memory(self):
        """(bytes held, bytes of released temporaries kept for reuse)."""
        held, pooled = C.c_int64(), C.c_int64()
        self.check(lib().epsx_device_memory(self.handle, C.byref(held), C.byref(pooled)))
        return held.value, pooled.value

    def ladder_timings(self):
        """Kernel seconds and launch counts of the last ladder run (classes of epsx_last_timings)."""
        s = np.zeros(8)
        n = np.zeros(8, dtype=np.int64)
        lib().epsx_last_timings(self.handle, ptr(s), ptr(n))
        return {"shift_gemm": s[0], "shift_gemm_launches": int(n[0]), "apply": s[1], "applications": int(n[1]),
                "bt_build": s[2], "bt_builds": int(n[2]), "instance_applications": int(n[3])}

    def timings(self):
        s = np.zeros(8)
        n = np.zeros(8, dtype=np.int64)
        lib().epsx_last_timings(self.handle, ptr(s), ptr(n))
        return {"fan_migdal": s[0], "kramers_kronig": s[1], "dos_mu": s[2], "other": s[3],
                "star_reduce": s[4], "star_gemm": s[5], "fan_migdal_launches": int(n[0]),
                "star_reduce_launches": int(n[4]), "star_gemm_launches": int(n[5]),
                "star_gemm_instances": int(n[6]), "star_gemm_kernel": int(n[7])}


_default = None


def default_context():
    """A process-wide context on device 0 (created on first use)."""
    global _default
    if _default is None:
        _default = Context(0)
    return _default
