"""Linearised operator for normal-mode analysis (imx_swm_jacobian,
swm_normal_mode_analysis src/swm.f90:738-976): the coloured finite
differences equal the reference's one-column-at-a-time differences, taken
with the oracle's tendency (bit-identical to the reference), bit for bit."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from imodel_amd.build import IMODEL
from imodel_amd.swm import Mesh, Swm
from oracle.oracle import ERAD_REF, Oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print("[gpu-case]", *a, flush=True)


def per_column(o, h, u, pert):
    """The reference's loop: one tendency per unknown."""
    nv, ne = len(h), len(u)
    m0, f0 = o.tendency(h, u)
    M = np.zeros((nv + ne, nv + ne))
    for i in range(nv + ne):
        hp, up = h.copy(), u.copy()
        if i < nv:
            hp[i] = h[i] + pert
        else:
            up[i - nv] = u[i - nv] + pert
        m, f = o.tendency(hp, up)
        M[:nv, i] = (m - m0) / pert
        M[nv:, i] = (f - f0) / pert
    return M


@pytest.mark.parametrize("glevel,state", [(1, "tc2"), (2, "spectral")])
def test_coloured_jacobian_equals_per_column(glevel, state):
    say("jacobian", glevel, state)
    m = Mesh.icos(glevel)
    h, u = m.tc2() if state == "tc2" else m.spectral(5)
    s = Swm(m, precision="fp64_ref", dt=900.0, erad=ERAD_REF)
    pert = float(np.mean(m.arrays["hx_areag"]))
    J, nev = s.jacobian(h, u)
    M = per_column(Oracle(m.arrays), h, u, pert)
    assert np.array_equal(J.toarray(), M)
    n = m.nv + m.ne
    assert nev < n / 3, (nev, n)
    assert J.nnz > 10 * n


def test_fortran_normal_mode_file():
    say("fortran normal modes g1")
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(os.path.join(td, "par"))
        src = os.path.join(ROOT, "par")
        mesh = open(os.path.join(src, "mesh.par")).read().splitlines()
        mesh[2] = "42"   # glevel 1: 162 unknowns, sparse file
        open(os.path.join(td, "par", "mesh.par"), "w").write("\n".join(mesh) + "\n")
        swm = open(os.path.join(src, "swm.par")).read().splitlines()
        swm[2] = "2 0"
        swm[6] = "900 0 0"
        open(os.path.join(td, "par", "swm.par"), "w").write("\n".join(swm) + "\n")
        gpu = open(os.path.join(src, "swm_gpu.par")).read().splitlines()
        gpu[1] = "fp64_ref"
        gpu[9] = "1"
        open(os.path.join(td, "par", "swm_gpu.par"), "w").write("\n".join(gpu) + "\n")
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        rows = np.loadtxt(os.path.join(td, "data", "swm_tc2_dt900_HTC_trsk10_areageo_linearmatrixsparse_icos_eqs_nopt_1.txt"))
    m = Mesh.icos(1)
    h, u = m.tc2()
    s = Swm(m, precision="fp64_ref", dt=900.0, erad=ERAD_REF)
    J, _ = s.jacobian(h, u)
    J = J.tocoo()
    keep = np.abs(J.data) > np.float32(np.finfo(np.float32).eps) / 100
    want = sorted(zip(J.row[keep] + 1, J.col[keep] + 1, J.data[keep]))
    got = sorted((int(a), int(b), c) for a, b, c in rows)
    assert len(got) == len(want)
    assert all(g[0] == w[0] and g[1] == w[1] and g[2] == w[2] for g, w in zip(got, want))
