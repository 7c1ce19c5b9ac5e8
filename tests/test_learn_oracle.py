"""CPU checks of the learned-correction restatement (oracle/learn_oracle.py)
and of the restriction operators (imodel_amd/learn.py): mimetic and
closed-form identities that a wrong index, sign or weight breaks."""
import numpy as np
import pytest

from imodel_amd.learn import restriction
from imodel_amd.swm import Mesh
from oracle import learn_oracle as lo
from oracle.oracle import ERAD_REF, Oracle


@pytest.fixture(scope="module")
def g3():
    return Mesh.icos(3)


def test_ring_signs_mimetic_identity(g3):
    """sum over the ring of c of s_k u_k l~_k plus ttgout*tnrccw u_e l~_e equals
    tnrccw A_c erad div_hx(u)_c; zero for the curl of a streamfunction."""
    a = g3.arrays
    st_e, st_s, st_c, geo = lo.stencils(a)
    dl = a["edhx_leng"]
    ev = a["ed_v"].reshape(-1, 2) - 1
    ved = a["v_ed"].reshape(-1, 6) - 1
    ttg = a["hx_ttgout"].reshape(-1, 6)
    tnr = a["hx_tnrccw"].reshape(-1, 6)
    nnb = a["v_nnb"]
    h, u = g3.spectral(3)
    o = Oracle(g3.full_arrays(), erad=ERAD_REF)
    o.tendency(h, u)
    divu = o.field("divu")
    # a generic field too: the identity holds for any u (div_hx from the oracle)
    u2 = np.random.default_rng(0).standard_normal(g3.ne)
    o.tendency(h, u2)
    divu2 = o.field("divu")
    A = a["hx_areag"]
    for uu, dv, tol in ((u, divu, 1e-9), (u2, divu2, 1e-9)):
        worst = 0.0
        for e in range(g3.ne):
            for q in range(2):
                c = ev[e, q]
                i = int(np.nonzero(ved[c, :nnb[c]] == e)[0][0])
                ks = range(1 + 5 * q, 6 + 5 * q)
                s = sum(st_s[e, k] * uu[st_e[e, k]] * dl[st_e[e, k]] for k in ks if st_e[e, k] >= 0)
                s += ttg[c, i] * tnr[c, i] * uu[e] * dl[e]
                worst = max(worst, abs(s - tnr[c, i] * A[c] * ERAD_REF * dv[c]) / np.abs(uu).max())
        assert worst < tol, worst
    assert np.abs(divu).max() * ERAD_REF < 1e-9 * np.abs(u).max()  # curl of psi: div-free


def test_stencil_padding_and_apexes(g3):
    a = g3.arrays
    st_e, st_s, st_c, geo = lo.stencils(a)
    # one padded ring slot and one padded ring cell per (edge, pentagon) pair
    assert int(np.sum(st_e < 0)) == 12 * 5
    assert int(np.sum(st_c < 0)) == 12 * 5
    p = a["v_p"].reshape(-1, 3)
    ev = a["ed_v"].reshape(-1, 2) - 1
    tv = a["tr_v"].reshape(-1, 3) - 1
    esh = a["ed_sh"].reshape(-1, 2) - 1
    for e in range(g3.ne):
        c1, c2, left, right = st_c[e, :4]
        assert (c1, c2) == tuple(ev[e])
        third = {int(v) for t in esh[e] for v in tv[t]} - {int(c1), int(c2)}
        assert third == {int(left), int(right)}
        n = np.cross(p[c1], p[c2])
        assert np.dot(n, p[left]) > 0 > np.dot(n, p[right])
        # ring cells are neighbours of c1 or c2, distinct from the first four
        rest = [int(x) for x in st_c[e, 4:] if x >= 0]
        assert len(set(rest)) == len(rest) and not set(rest) & {int(c1), int(c2), int(left), int(right)}
    assert np.allclose(geo[:, 0] ** 2 + geo[:, 1] ** 2, 1.0)


def test_div_matrix_conserves_mass_and_matches_div_hx(g3):
    a = g3.arrays
    D = lo.div_matrix(a)
    A = a["hx_areag"]
    F = np.random.default_rng(1).standard_normal(g3.ne)
    DF = D @ F
    assert abs(np.sum(A * DF)) < 1e-13 * np.sum(A * np.abs(DF))
    o = Oracle(g3.full_arrays(), erad=ERAD_REF)
    h, _ = g3.spectral(1)
    o.tendency(h, F)
    kappa = A.mean() / a["edhx_leng"].mean()
    assert np.allclose(DF, -kappa * ERAD_REF * o.field("divu"), rtol=1e-12, atol=1e-12 * np.abs(DF).max())


def test_features_layout(g3):
    a = g3.arrays
    st = lo.stencils(a)
    h0, u0 = g3.spectral(2)
    h1, u1 = h0 + 1.0, u0 * 1.5
    raw = lo.features(st, h0, u0, h1, u1, 3000.0)
    assert raw.shape == (g3.ne, 48)
    assert np.allclose(raw[:, 0], u0)                     # the edge itself, sign +1
    assert np.allclose(raw[:, 11], 0.5 * u0)              # its increment
    ev = a["ed_v"].reshape(-1, 2) - 1
    assert np.allclose(raw[:, 22], h0[ev[:, 0]] - 3000.0)
    assert np.allclose(raw[:, 32], 1.0)                   # dh of c1
    pad = st[0] < 0
    assert np.all(raw[:, :11][pad] == 0) and np.all(raw[:, 11:22][pad] == 0)


def test_correct_limiter():
    rng = np.random.default_rng(3)
    from scipy.sparse import identity
    h, u = rng.standard_normal(5), rng.standard_normal(7)
    Y = rng.standard_normal((7, 2))
    D = identity(7, format="csr")[:5]
    h1, u1, r = lo.correct(h, u, Y, D, 2.0, 4.0, 0.0)
    assert h1 is h and u1 is u and r > 0
    h2, u2, _ = lo.correct(h, u, Y, D, 2.0, 4.0, 1e9)
    assert np.allclose(u2 - u, Y[:, 0] / 2.0) and np.allclose(h2 - h, Y[:5, 1] / 4.0)
    # above the bound the correction is scaled to RMS = bound
    h3, u3, _ = lo.correct(h, u, Y, D, 1.0, 1.0, 0.25 * r)
    assert np.allclose(u3 - u, 0.25 * Y[:, 0]) and np.allclose(h3 - h, 0.25 * Y[:5, 1])
    dh, du = h3 - h, u3 - u
    assert np.isclose(np.sqrt(np.mean(du ** 2) + np.mean(dh ** 2)), 0.25 * r)
    Yn = Y.copy()
    Yn[0, 0] = np.nan
    h4, u4, _ = lo.correct(h, u, Yn, D, 1.0, 1.0, 1.0)
    assert h4 is h and u4 is u


@pytest.fixture(scope="module")
def levels():
    return {g: Mesh.icos(g) for g in (2, 3, 4)}


def test_restriction_composes_and_preserves_constants(levels):
    R42h, R42u = restriction(levels[4], levels[2], levels)
    R43h, R43u = restriction(levels[4], levels[3], levels)
    R32h, R32u = restriction(levels[3], levels[2], levels)
    assert abs(R42h - R32h @ R43h).max() == 0.0
    assert abs(R42u - R32u @ R43u).max() == 0.0
    assert np.allclose(R42h @ np.ones(levels[4].nv), 1.0, rtol=0, atol=1e-14)
    assert (R42h.data > 0).all()
    # u: |weights| of a coarse edge sum to 1 over 4 fine edges
    assert np.allclose(abs(R42u).sum(axis=1), 1.0)
    assert np.all(np.diff(R42u.indptr) == 4)


def test_restriction_solid_body_exact(levels):
    """TC2's velocity is solid-body rotation, whose tangential component is
    constant along any great circle: the edge restriction reproduces the
    coarse field exactly; a wrong sign or a wrong fine edge breaks this."""
    Rh, Ru = restriction(levels[4], levels[2], levels)
    hf, uf = levels[4].tc2()
    hc, uc = levels[2].tc2()
    assert np.abs(Ru @ uf - uc).max() < 1e-12 * np.abs(uc).max()
    # h: cell means of a smooth field, second-order close to point values
    assert np.abs(Rh @ hf - hc).max() < 0.03 * np.abs(hc).max()


def test_restriction_rejects_non_nested():
    with pytest.raises(ValueError):
        restriction(Mesh.icos(2), Mesh.icos(3))
