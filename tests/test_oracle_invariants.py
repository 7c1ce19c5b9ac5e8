"""Invariants of the discretisation that pin the oracle beyond the golden
vectors (closed forms and mimetic identities of the TRiSK C-grid scheme,
cf. the checks of src/swm.f90:625-718)."""
import numpy as np

from imodel_amd.swm import Mesh
from oracle.oracle import Oracle


def _mesh(g):
    return Mesh.icos(g)


def test_mass_conserved_fp64():
    m = _mesh(4)
    o = Oracle(m.arrays)
    h, u = m.spectral(11)
    A = m.arrays["hx_areag"]
    m0 = np.sum(A * h)
    h2, u2 = o.step(h, u, 450.0, 10)
    assert abs(np.sum(A * h2) - m0) / m0 < 1e-13


def test_coriolis_energy_neutral():
    """TRiSK: sum_e uh_e uhq_perp_e l_e d_e = 0 (src/swm.f90:704-718)."""
    m = _mesh(4)
    o = Oracle(m.arrays)
    h, u = m.spectral(5)
    o.tendency(h, u)
    uh, uhq = o.field("uh"), o.field("uhq_perp")
    le, de = m.arrays["ed_leng"], m.arrays["edhx_leng"]
    s = np.sum(uh * uhq * le * de)
    scale = np.sum(np.abs(uh * uhq * le * de))
    assert abs(s) / scale < 1e-13


def test_curl_grad_zero():
    """rel_vort_tr of a discrete gradient vanishes (src/swm.f90:667-700)."""
    m = _mesh(3)
    o = Oracle(m.arrays)
    phi = np.random.default_rng(0).standard_normal(m.nv)
    ev = m.arrays["ed_v"].reshape(-1, 2) - 1
    u = (phi[ev[:, 1]] - phi[ev[:, 0]]) / m.arrays["ed_leng"]
    h = np.full(m.nv, 1000.0)
    o.tendency(h, u)
    z = o.field("zeta")
    assert np.max(np.abs(z)) < 1e-12 * np.max(np.abs(u)) / 6.37122e6


def test_streamfunction_state_nondivergent():
    """The seeded spectral velocity is a discrete curl: div u = 0 to round-off."""
    m = _mesh(4)
    o = Oracle(m.arrays)
    h, u = m.spectral(2)
    o.tendency(h, u)
    divu = o.field("divu")
    # scale: |u| / (a * dx)
    scale = np.max(np.abs(u)) / 6.37122e6 / np.min(m.arrays["edhx_leng"])
    assert np.max(np.abs(divu)) < 1e-12 * scale


def test_tc2_steady_and_converging():
    """Williamson TC2 is a steady state: after one day the deviation is the
    scheme's truncation error, which shrinks with refinement (g4 -> g5); a
    wrong sign in any term would blow it up instead."""
    errs = []
    for g in (4, 5):
        m = _mesh(g)
        o = Oracle(m.arrays)
        h, u = m.tc2()
        n = 96 * 2 ** (g - 4)
        h2, u2 = o.step(h, u, 86400.0 / n, n)
        errs.append(np.max(np.abs(h2 - h)) / np.max(h))
    assert errs[0] < 5e-3
    assert errs[1] < 0.6 * errs[0]


def test_potential_enstrophy_rest_state():
    """Fluid at rest, h = H: q_t = f_t/H, so sum_t 1/2 A_t a^2 h_t q_t^2
    = 1/2 a^2 (2 Omega)^2 / H * sum_t A_t z_t^2, and the triangle quadrature
    of z^2 over the icosahedral sphere is exact: 4 pi / 3."""
    H, a, Om = 1000.0, 6.37122e6, 7.292e-5
    exact = 0.5 * a * a * (2 * Om) ** 2 / H * 4 * np.pi / 3
    for g in (3, 4):
        m = _mesh(g)
        o = Oracle(m.arrays)
        d = o.diagnostics(np.full(m.nv, H), np.zeros(m.ne))
        assert abs(d[5] / exact - 1) < 1e-12
