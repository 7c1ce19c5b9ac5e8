"""imx_swm_diagnostics_dev: the diagnostics of a handle written into the
caller's device buffer with the reduction partials in the library's static
device array (no allocation), against imx_swm_diagnostics on the same
handle: two calls in a row, the second on a smaller mesh, with and without
topography and in both fused precisions and the reference order."""
import ctypes
import os

import numpy as np
import pytest

from imodel_amd import capi
from imodel_amd.swm import Mesh, Swm

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def hip_runtime():
    """The HIP runtime libimx itself links (one runtime in the process), for
    the caller's device buffer."""
    capi.lib()
    path = "/opt/rocm/lib/libamdhip64.so.7"
    H = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return H


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def test_diagnostics_dev_matches_diagnostics():
    L = capi.lib()
    L.imx_swm_diagnostics_dev.restype = ctypes.c_int
    L.imx_swm_diagnostics_dev.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    H = hip_runtime()
    nb = capi.NDIAG * 8
    out = ctypes.c_void_p()  # the caller's output, allocated by the caller
    assert H.hipMalloc(ctypes.byref(out), nb) == 0
    try:
        for g, prec, seed, topo in ((5, "fp32", 1, True), (3, "fp64", 2, False), (2, "fp64_ref", 3, True)):
            say("diagnostics_dev", g, prec, seed, topo)
            m = Mesh.icos(g)
            h, u = m.spectral(seed)
            with Swm(m, precision=prec, dt=900.0) as s:
                if topo:
                    s.set_topography(m.testcase(5)[2])
                s.set_state(h, u)
                s.step(2)
                ref = s.diagnostics()
                nan = np.full(capi.NDIAG, np.nan)
                assert H.hipMemcpy(out, nan.ctypes.data, nb, H2D) == 0
                capi.check(L.imx_swm_diagnostics_dev(s._h, out), "imx_swm_diagnostics_dev")
                assert H.hipDeviceSynchronize() == 0
                got = np.zeros(capi.NDIAG)
                assert H.hipMemcpy(got.ctypes.data, out, nb, D2H) == 0
            want = np.array([ref[k] for k in capi.DIAG_NAMES])
            # the same partial sums in the same order; the finish (divisions,
            # a square root) runs on the device instead of the host
            assert np.all(np.isfinite(got)), got
            assert np.allclose(got, want, rtol=1e-15, atol=0), (got, want)
    finally:
        H.hipFree(out)
