"""imx_swm_diagnostics_dev: the diagnostics of a handle written into the
caller's device buffer with the reduction partials in the library's static
device array (no allocation), against imx_swm_diagnostics on the same
handle: two calls in a row, the second on a smaller mesh, with and without
topography and in both fused precisions and the reference order."""
import ctypes

import numpy as np
import pytest

from imodel_amd import capi
from imodel_amd.swm import Mesh, Swm

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def test_diagnostics_dev_matches_diagnostics():
    import torch

    L = capi.lib()
    L.imx_swm_diagnostics_dev.restype = ctypes.c_int
    L.imx_swm_diagnostics_dev.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    out = torch.zeros(capi.NDIAG, dtype=torch.float64, device="cuda")  # the caller's output
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
            out.fill_(float("nan"))
            torch.cuda.synchronize()
            capi.check(L.imx_swm_diagnostics_dev(s._h, ctypes.c_void_p(out.data_ptr())), "imx_swm_diagnostics_dev")
            torch.cuda.synchronize()
            got = out.cpu().numpy()
        want = np.array([ref[k] for k in capi.DIAG_NAMES])
        # the same partial sums in the same order; the finish (divisions, a
        # square root) runs on the device instead of the host
        assert np.all(np.isfinite(got)), got
        assert np.allclose(got, want, rtol=1e-15, atol=0), (got, want)
