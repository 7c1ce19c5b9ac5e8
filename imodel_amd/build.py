"""In-tree build of libimx (HIP for gfx950 + host C++) and the Fortran host.

Everything is compiled every time ``build()`` is called; outputs stay inside
the repository (``imodel_amd/libimx.so``, ``imodel_amd/_build/``,
``imodel_amd/fortran/imodel``) so that they travel with the source tree.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libimx.so")
IMODEL = os.path.join(FSRC, "imodel")
# a test program against the swm call surface (tests/fortran), built beside it
CALLER_SRC = os.path.join(os.path.dirname(HERE), "tests", "fortran", "swm_caller.f90")
CALLER = os.path.join(os.path.dirname(HERE), "tests", "fortran", "swm_caller")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC_IMX", "/opt/rocm/bin/amdflang")
ARCH = "gfx950"

# -ffp-contract=off everywhere: FP64 results follow the written expression
# order (the REF kernels rely on it); packed FP32 ops are disabled for gfx950.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
             "-Wno-unused-function"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]

HOST_SRCS = ["imx_mesh.cpp", "imx_trmesh.cpp", "imx_ic.cpp", "imx_sfc.cpp", "imx_gridio.cpp", "imx_plot.cpp", "imx_jacobian.cpp",
             "imx_lte.cpp"]
HIP_SRCS = ["imx_swm.hip", "imx_mlp.hip", "imx_learn.hip"]
# the Fortran host's modules, in dependency order
FORTRAN_SRCS = ("datastruct.f90", "imx_c.f90", "swm_par.f90", "smeshpack.f90", "swm_data.f90", "swm.f90")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_lib():
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for s in HOST_SRCS:
        o = os.path.join(BUILD, s.replace(".cpp", ".o"))
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRCS:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """The Fortran host (imodel driver + swm module over the C ABI), and the
    test program written against the swm call surface."""
    if not os.path.isdir(FSRC):
        return None
    srcs = [os.path.join(FSRC, f) for f in FORTRAN_SRCS]
    if not all(os.path.exists(s) for s in srcs):
        return None
    mod = os.path.join(BUILD, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in srcs:  # in dependency order: each module's .mod before its users
        o = os.path.join(BUILD, os.path.basename(src).replace(".f90", ".fo"))
        _run([FC, "-O2", "-module-dir", mod, "-I", mod, "-c", src, "-o", o])
        objs.append(o)
    link = [f"-L{HERE}", "-limx", f"-Wl,-rpath,{HERE}"]
    _run([FC, "-O2", "-I", mod, os.path.join(FSRC, "imodel.f90"), *objs, "-o", IMODEL, *link,
          "-Wl,-rpath,$ORIGIN/.."])
    if os.path.exists(CALLER_SRC):
        _run([FC, "-O2", "-I", mod, CALLER_SRC, *objs, "-o", CALLER, *link,
              "-Wl,-rpath,$ORIGIN/../../imodel_amd"])
    return IMODEL


def build():
    lib = build_lib()
    exe = build_fortran()
    return lib, exe


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
