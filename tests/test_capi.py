"""libimx loads on a CPU-only host and exports every symbol include/imx_swm.h
declares (no compute calls here)."""
import ctypes
import os
import re

from imodel_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            src = open(os.path.join(ROOT, "include", h)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"\b(imx_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_exports_every_declared_symbol():
    L = ctypes.CDLL(capi.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_every_export_is_declared():
    """The other direction: no imx_ entry point is exported without a
    declaration in include/."""
    import shutil
    import subprocess
    nm = shutil.which("nm")
    if nm is None:
        import pytest
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", capi.LIB_PATH], check=True, stdout=subprocess.PIPE,
                         text=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines() if ln.split() and
                       ln.split()[-1].startswith("imx_")})
    undeclared = sorted(set(exported) - set(declared()))
    assert not undeclared, undeclared


def test_error_paths_without_gpu():
    L = capi.lib()
    h = ctypes.c_void_p()
    assert L.imx_mesh_build_icos(99, 0, ctypes.byref(h)) == -1
    assert b"glevel" in L.imx_last_error()
    assert L.imx_swm_create(None, None, None) == -1
