"""Build libepsx.so (gfx950 HIP kernels + the C ABI) in-tree with hipcc.

    python epspectral.jl_amd/build.py            # the product library
    python epspectral.jl_amd/build.py --check    # build and import the Python package

The library lands in epspectral.jl_amd/epsx/libepsx.so, next to the Python mirror that loads
it.  Every HIP translation unit is compiled for gfx950 only, with packed FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "epsx", "libepsx.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = [
    "epsx_context.cpp",
    "epsx_grids.cpp",
    "epsx_kk.hip", "epsx_gemm.hip", "epsx_net.hip",
    "epsx_fm.hip",
    "epsx_fm_star.hip",
    "epsx_scgd0.hip",
    "epsx_ladder.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + INCLUDE, "-I" + CSRC,
]


def _compile(src):
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    subprocess.run(cmd, check=True)
    return obj


def build(jobs=None):
    """Compile every source (in parallel) and link libepsx.so.  Always rebuilds."""
    from concurrent.futures import ThreadPoolExecutor

    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    # rocBLAS serves the ladder's plain shift-table GEMM and hipBLASLt the star method's FP64
    # product (the ROCm install's libraries, found at run time through the rpath)
    libs = ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocblas", "-lhipblaslt"]
    subprocess.run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + libs, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    path = build()
    print("built", path)
    if "--check" in sys.argv:
        sys.path.insert(0, HERE)
        import epsx  # noqa: F401

        print("imported epsx from", epsx.__file__)
