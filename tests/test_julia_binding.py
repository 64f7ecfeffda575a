"""The Julia binding (epspectral.jl_amd/julia/EPSpectralMI355X.jl) against the C ABI (CPU only).

Julia is not installed here, so the binding is checked as text: every struct mirror has the fields,
order and C types of the corresponding ctypes structure (whose sizes tests/test_abi.py pins against
the header), every `ccall` names an entry point include/epsx.h declares, and the context's
finalizer is attached to the mutable backend object itself.
"""
import ctypes
import os
import re

from epsx import _lib
from test_abi import declared

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = open(os.path.join(ROOT, "epspectral.jl_amd", "julia", "EPSpectralMI355X.jl")).read()

JL_TYPES = {"Int64": ctypes.c_int64, "Int32": ctypes.c_int32, "Float64": ctypes.c_double,
            "UInt64": ctypes.c_uint64}
PAIRS = {"Scgd0Problem": _lib.Scgd0Problem, "Scgd0Params": _lib.Scgd0Params, "Scgd0Outputs": _lib.Scgd0Outputs,
         "LadderProblem": _lib.LadderProblem, "LadderParams": _lib.LadderParams,
         "LadderOutputs": _lib.LadderOutputs, "NetConfig": _lib.NetConfig, "NetTrainParams": _lib.NetTrainParams}


def jl_struct(name):
    m = re.search(r"(?:mutable\s+)?struct\s+" + name + r"\b(.*?)\nend", JL, flags=re.S)
    assert m, name
    fields = []
    for decl in re.findall(r"(\w+)::([\w{}]+)", m.group(1)):
        fields.append(decl)
    return fields


def test_struct_mirrors_match_the_abi():
    for jname, cstruct in PAIRS.items():
        jf = jl_struct(jname)
        cf = cstruct._fields_
        assert [n for n, _ in jf] == [n for n, _ in cf], jname
        for (n, jt), (_, ct) in zip(jf, cf):
            if jt.startswith("Ptr"):
                assert ct is ctypes.c_void_p, (jname, n)
            else:
                assert JL_TYPES[jt] is ct, (jname, n, jt, ct)


def test_every_ccall_is_a_declared_entry_point():
    called = set(re.findall(r"ccall\(\(:(epsx_\w+),", JL))
    assert called <= set(declared())
    for name in ("epsx_scgd0_run", "epsx_ladder_run", "epsx_ladder_set_problem", "epsx_net_forward",
                 "epsx_net_train", "epsx_close"):
        assert name in called, name


def test_finalizer_on_the_backend_object():
    assert re.search(r"mutable struct MI355XBackend", JL)
    assert re.search(r"finalizer\(close!, b\)", JL)
    assert "Ref(b)" not in JL
    # the reference entry points the binding mirrors
    for fn in ("loop_scgd0", "loop_ladder", "run_scgd0_mi355x", "run_scgd0_ladder_mi355x", "close!"):
        assert re.search(r"function " + re.escape(fn) + r"\(", JL), fn
